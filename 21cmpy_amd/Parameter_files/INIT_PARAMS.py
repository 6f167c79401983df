"""Run configuration with the reference's names and default values.

The reference keeps its configuration as module globals in
Parameter_files/INIT_PARAMS.py (star-imported by every stage); the same names
are kept here so existing scripts find them.  Launch-geometry globals of the
reference (block_size, grid_size, HII_grid_size) have no meaning for this
implementation and are not defined.
"""
import numpy as np

IO_DIR = "."                                   # output root; Boxes/ lives below it

DIM = np.int32(512)                            # high-resolution box side (cells)
HII_DIM = np.int32(128)                        # ionization box side (cells)
PIXEL_FACTOR = np.float32(DIM) / np.float32(HII_DIM)
MASS_FACTOR = np.float32(PIXEL_FACTOR ** 3.0)
BOX_LEN = np.float32(128)                      # comoving Mpc
DELTA_K = np.float32(2 * np.pi / BOX_LEN)
VOLUME = BOX_LEN * BOX_LEN * BOX_LEN
HII_TOT_NUM_PIXELS = HII_DIM * HII_DIM * HII_DIM
HII_shape = (int(HII_DIM),) * 3

INITIAL_REDSHIFT = 300.0
EVOLVE_DENSITY_LINEARLY = False
ION_Tvir_MIN = 1.0e4
ZETA = 40.0
R_BUBBLE_MAX = 30.0                            # Mpc
R_BUBBLE_MIN = 0.5                             # Mpc
FIND_BUBBLE_ALGORITHM = 0                      # 0 centre pixel, 1 sphere paint
DELTA_R_FACTOR = 1.1

# 0 real-space top-hat, 1 k-space top-hat, 2 Gaussian
HII_FILTER = 1
L_FACTOR = np.float32(0.620350491)             # (4 pi/3)^(-1/3)

PLANCK13 = {
    "baryonic_effects": False,
    "omega_k_0": 0,
    "omega_M_0": 0.315,
    "omega_b_0": 0.0487,
    "n": 0.96,
    "N_nu": 0,
    "omega_lambda_0": 0.685,
    "omega_n_0": 0.0,
    "sigma_8": 0.829,
    "h": 0.673,
}
COSMO = PLANCK13

ZERO = np.int32(0)
ONE = np.int32(1)
TWO = np.int32(2)

ZETA_X = 2.0e56
DIMENSIONAL_T_POWER_SPEC = False
T_USE_VELOCITIES = False
NUM_BINS = 100
USE_TS_IN_21CM = 0
