lx = 40e-9
ly = 40e-9
lz = 40e-9
default_material(0)
local r = 7e-9
for i, c in ipairs({{12e-9, 13e-9, 14e-9}, {27e-9, 26e-9, 25e-9}}) do
    add_sphere(c[1], c[2], c[3], r + i*1e-9, 1)
end
