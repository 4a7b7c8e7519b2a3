-- Normal incidence with the automatic run length (energy criterion, checked
-- every 10 steps, stop below 1e-3 of the peak energy).
st = Structure("slab_structure.lua")
st:parameter("h_top", 10e-9)

f = MODE("fdtd_normal")
f:prefix("auto")
f:polarization(30)          -- a number: "mix" at 30 degrees
f:structure(st)
f:Dxyz(5e-9)
f:auto_tsteps(4000, 10, 1e-3)
f:spectrum(400e-9, 800e-9, 12)
f:pml_zp(8, 25, 1, 0.2)
f:pml_zm(8, 25, 1, 0.2)
f:material(1, Material(1.5))
f:material(2, "mat_lorentz.lua")
f:compute()
