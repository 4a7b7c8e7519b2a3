-- A dielectric slab with a Lorentz layer on top, thicknesses as parameters.
declare_parameter("h_slab", 60e-9)
declare_parameter("h_top", 20e-9)

lx = 40e-9
ly = 30e-9
lz = 200e-9

default_material(0)
add_layer("Z", 60e-9, 60e-9 + h_slab, 1)
add_layer("z", 60e-9 + h_slab, 60e-9 + h_slab + h_top, 2)
