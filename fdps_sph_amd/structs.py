"""numpy dtypes byte-compatible with the reference's boundary types
(src/class.h:36-86, 233-453) and FDPS's PS::SPJMonopole; the same layouts are
declared as C structs in include/fdps_sph_gpu.h."""
import numpy as np

f8 = "<f8"
i8 = "<i8"

EPI_DENS = np.dtype([("pos", f8, 3), ("mass", f8), ("smth", f8), ("id", i8)])
EPJ_DENS = np.dtype([("mass", f8), ("pos", f8, 3), ("smth", f8)])
RES_DENS = np.dtype([("dens", f8), ("smth", f8)])
EPI_DRVT = np.dtype([("pos", f8, 3), ("vel", f8, 3), ("smth", f8), ("dens", f8)])
EPJ_DRVT = np.dtype([("mass", f8), ("pos", f8, 3), ("vel", f8, 3), ("smth", f8)])
RES_DRVT = np.dtype([("div_v", f8), ("rot_v", f8, 3), ("grad_smth", f8)])
EPI_HYDRO = np.dtype([("pos", f8, 3), ("vel", f8, 3), ("smth", f8), ("dens", f8), ("pres", f8), ("snds", f8),
                      ("grad_smth", f8), ("Bal", f8), ("AVa", f8), ("id", i8)])
EPJ_HYDRO = np.dtype([("pos", f8, 3), ("vel", f8, 3), ("dens", f8), ("mass", f8), ("smth", f8), ("pres", f8),
                      ("grad_smth", f8), ("snds", f8), ("Bal", f8), ("AVa", f8), ("id", i8)])
RES_HYDRO = np.dtype([("acc", f8, 3), ("eng_dot", f8), ("dt", f8)])
EPI_GRAV = np.dtype([("pos", f8, 3), ("eps2", f8), ("id", i8)])
EPJ_GRAV = np.dtype([("pos", f8, 3), ("mass", f8), ("id", i8)])
SPJ_MONOPOLE = np.dtype([("mass", f8), ("pos", f8, 3)])
RES_GRAV = np.dtype([("acc", f8, 3), ("pot", f8), ("dt", f8)])

SIZES = {"EPI_DENS": 48, "EPJ_DENS": 40, "RES_DENS": 16, "EPI_DRVT": 64, "EPJ_DRVT": 64, "RES_DRVT": 40,
         "EPI_HYDRO": 112, "EPJ_HYDRO": 120, "RES_HYDRO": 40, "EPI_GRAV": 40, "EPJ_GRAV": 40,
         "SPJ_MONOPOLE": 32, "RES_GRAV": 40}
