"""Stage programs with the reference's module names: initialize,
perturb_field, find_HII_bubbles, delta_T (and the Drives redshift driver)."""
