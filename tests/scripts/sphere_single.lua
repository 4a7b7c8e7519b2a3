-- Single particle: two spheres of a dispersive metal in a dielectric host.
st = Structure("spheres_structure.lua")

box = create_sensor("box_spectral_poynting")
box:location(0, 40e-9, 0, 40e-9, 0, 40e-9)
box:spectrum(400e-9, 800e-9, 10)
box:name("box")

f = MODE("fdtd_single_particle")
f:prefix("sph_")
f:polarization("TM")
f:structure(st)
f:Dxyz(2e-9)
f:N_tsteps(150)
f:spectrum(400e-9, 800e-9, 10)
for _, face in ipairs({"xm", "xp", "ym", "yp", "zm", "zp"}) do
    f["pml_" .. face](f, 5, 25, 1, 0.2)
end
f:material(0, const_material(1.33))
f:material(1, "mat_metal.lua")
f:register_sensor(box)
f:compute()
