-- Normal incidence on the slab, TE then TM, second run with a thicker slab.
glass = Material()
glass:refractive_index(1.5)
glass:name("glass")

st = Structure("slab_structure.lua")

local function run(pol, pfx)
    local f = MODE("fdtd_normal")
    f:prefix(pfx)
    f:polarization(pol)
    f:structure(st)
    f:Dxyz(5e-9)
    f:N_tsteps(400)
    f:spectrum(400e-9, 800e-9, 12)
    f:pml_zp(8, 25, 1, 0.2)
    f:pml_zm(8, 25, 1, 0.2)
    f:tapering(50)
    f:material(0, Material(1.0))
    f:material(1, glass)
    f:material(2, "mat_lorentz.lua")
    f:compute()
    return f
end

a = run("TE", "slab_" .. "TE")
b = run("TM", "slab_TM_")
st:parameter("h_slab", 80e-9)
c = run("te", "slab_thick")
if a:Nz() ~= 40 or c:Lz() ~= 200e-9 then fail_test() end
