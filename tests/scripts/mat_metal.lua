-- Test material: a Drude pole and two critical points.
name("test metal")
eV = 1.0/6.58211951440e-16
index_infty(math.sqrt(1.1))
add_drude(9.0*eV, 0.07*eV)
for k = 1, 2 do
    add_crit_point(0.8*k, (2.0 + k)*eV, -0.5*k, 0.5*eV)
end
