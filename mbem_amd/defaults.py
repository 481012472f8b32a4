"""The numeric constants the dense hot path reads (the used subset of mbem's
src/mbem/defaults.py, same names and values; cited line by line)."""

SOLUTION_RTOL = 1e-6               # defaults.py:10
GMRES_RTOL = 1e-8                  # defaults.py:11
GMRES_RESTART = 200                # defaults.py:12
GMRES_MAXITER = 600                # defaults.py:13
GMRES_STAGNATION_WINDOW = 100      # defaults.py:16
GMRES_STAGNATION_FACTOR = 10.0     # defaults.py:17
PRECOND_GROUPING = "patch"         # defaults.py:63
COND_WARN_THRESHOLD = 1e10         # defaults.py:118
CLUSTER_MIN_LEAF = 96              # defaults.py:132
MAX_DENSE_PRECOND_DOF = 30_000     # defaults.py:645
PRECOND_DENSE_RAM_FRACTION = 0.06  # defaults.py:653
PRECOND_RUNG_ABOVE_DENSE = "block_jacobi"  # defaults.py:695
PRECOND_BJ_CHUNK_DOF = 9000        # defaults.py:712
PRECOND_REUSE_MAX_STEP = 0.6931471805599453   # defaults.py:78 (= log 2)
GCRO_RECYCLE_DIM = 30              # defaults.py:94
GCRO_RECYCLE_DEFAULT = False       # defaults.py:99
EPS_OVER_H = 0.1                   # defaults.py:798
FAULT_EPS_OVER_H = 0.07            # defaults.py:804
HALF_JUMP_MAX_EPS_OVER_H = 0.5     # defaults.py:808
NEAR_BOUNDARY_H_RATIO = 0.5        # defaults.py:824
EIGEN_NEAR_RADIUS_EPS = 20.0       # defaults.py:846
MIN_TRIANGLE_HEIGHT_OVER_L = 1e-6  # defaults.py:946
COLLOCATION_SHRINK_BY_ORDER = {0: 0.0, 1: 0.5, 2: 0.5}  # defaults.py:957
COLLOCATION_FAULT_CLEARANCE_EPS = 1.0  # defaults.py:963
ASSEMBLY_SOURCE_CHUNK = 256        # defaults.py:981
DENSE_LU = "double"                # defaults.py:1243
LU_REFINE_TOL = 1.0e-10            # defaults.py:1250
LU_REFINE_PASSES = 4               # defaults.py:1255
LU_REFINE_RATE = 1.0e-2            # defaults.py:1256
LAPLACE_POINTS_PER_DECADE = 3      # defaults.py:1035
LAPLACE_PAD_DECADES = 0.5          # defaults.py:1040
PRONY_RCOND = 1.0e-13              # defaults.py:1067
LAPLACE_TAIL_DECADES = 12.0        # defaults.py:1062
EXTENDED_BURGERS_PER_DECADE = 4    # defaults.py:1082
STEPPER_TAIL_DECADES = 1.0         # defaults.py:1117
NODAL_D_STAR = 10.0                # defaults.py:916
NODAL_FAR_GAUSS_N = 12             # defaults.py:917
NODAL_FAR_GAUSS_N_DISTANT = 8      # defaults.py:918
NODAL_D_STAR_DISTANT = 40.0        # defaults.py:919
NODAL_SERIES_U_OVER_RHO = 2.0      # defaults.py:925
NODAL_SERIES_TERMS = 32            # defaults.py:926
NODAL_SMALL_U_OVER_RHO = 0.5       # defaults.py:927
NODAL_MIN_HEIGHT_OVER_L = 1e-3     # defaults.py:932
