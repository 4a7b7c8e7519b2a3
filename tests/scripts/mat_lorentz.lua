-- Test material: one Lorentz pole (A, O, G) on eps_inf 2.
name("test lorentz")
description('one Lorentz pole\nfor the front-end tests')
ih = 1.0/6.58211951440e-16
epsilon_infty(2)
add_lorentz(1.2, 3.1*ih, 0.25*ih)
validity_range(400e-9, 800e-9)
