lx = 100e-9
ly = 90e-9
lz = 80e-9
default_material(0)
add_block(30e-9, 50e-9, 25e-9, 60e-9, 35e-9, 70e-9, 1)
