"""ctypes binding of liblcs_mi355x.so (the C ABI declared in include/lcs_mi355x.h).

The shared library is built in-tree by ``make -C lcs.jl_amd`` (or
``__graft_entry__.build()``).  There is no fallback: if the library is missing
or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liblcs_mi355x.so")

LCS_OK = 0
LCS_ERR_INVALID = -1
LCS_ERR_HIP = -2
LCS_ERR_CAPACITY = -3
LCS_ERR_STATE = -4

LCS_F64 = 0
LCS_F32 = 1

FIELDS = {"U": 0, "V": 1, "W": 2, "U2": 3, "V2": 4, "W2": 5, "P": 6, "dU": 7, "dV": 8, "dW": 9}
# Tuple(::Properties) order, src/particles/buffer.jl:32-35
PROPS = ["ids", "diams", "x", "y", "z", "x2", "y2", "z2", "u", "v", "w", "u2", "v2", "w2", "dudt", "dvdt", "dwdt"]
PROP_ID = {n: i for i, n in enumerate(PROPS)}

COLORING = {"red-black-fast": 0, "red-black": 1}
DRAG = {"linear": 0, "nonlinear": 1}
FORCING = {"no": 0, "linear-power": 1, "rc-energy": 2}


class LCSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class LCSArgumentError(LCSError, ValueError):
    pass


class BoxDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("halo", C.c_int), ("dtype", C.c_int),
                ("U0_m_s", C.c_double), ("L0_m", C.c_double), ("NU_m2_s", C.c_double)]


class HsmacParams(C.Structure):
    _fields_ = [("time_blocking", C.c_int), ("overlap", C.c_int), ("coloring", C.c_int), ("omega", C.c_double),
                ("epsp", C.c_double), ("itrmin", C.c_int), ("itrmax", C.c_int)]


class ParticleDesc(C.Structure):
    _fields_ = [("nvalid", C.c_int64), ("ntotal", C.c_int64), ("gravity", C.c_double),
                ("density_ratio", C.c_double), ("drag", C.c_int), ("grid_ratio", C.c_int)]


class Forcing(C.Structure):
    _fields_ = [("kind", C.c_int), ("power", C.c_double), ("energy", C.c_double), ("kmin", C.c_int),
                ("kmax", C.c_int), ("coarse_size", C.c_int)]


class StepConfig(C.Structure):
    _fields_ = [("cfl", C.c_double), ("dtmax", C.c_double), ("hsmac", HsmacParams), ("forcing", Forcing),
                ("localize_interval", C.c_int)]


class StepLog(C.Structure):
    _fields_ = [("step", C.c_int64), ("t", C.c_double), ("dt", C.c_double), ("itr1", C.c_int), ("itr2", C.c_int),
                ("div1", C.c_double), ("div2", C.c_double), ("p", C.c_double), ("e10", C.c_double)]


class Request(C.Structure):
    _fields_ = [("n", C.c_int), ("dtype", C.c_int), ("particles", C.c_int64), ("st_min", C.c_double),
                ("st_max", C.c_double), ("forcing", C.c_int), ("forcing_param", C.c_double), ("steps", C.c_int),
                ("seed", C.c_uint64), ("k0", C.c_double), ("Re", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_size_t), ("max_live", C.c_int), ("cfl", C.c_double), ("dtmax", C.c_double),
                ("hsmac", HsmacParams), ("localize_interval", C.c_int)]


FLOW_STAT_FIELDS = ["u_mean", "u_max", "u_min", "v_mean", "v_max", "v_min", "w_mean", "w_max", "w_min", "k", "eps",
                    "tau_eta", "l_mu", "l_eta", "Re_lambda", "urms", "k_eta", "skew", "flat", "L", "C_eps"]


class ParticleStatParams(C.Structure):
    _fields_ = [("contact_tolerance", C.c_double), ("bin_size", C.c_int), ("min_radius", C.c_double),
                ("search_cell_size", C.c_int), ("max_wavenumber", C.c_int), ("nsamples", C.c_int)]


PSTAT_SUMMARIES = [f"{c}_{k}" for c in ("x", "y", "z", "u", "v", "w", "diam") for k in ("mean", "max", "min")]


class ParticleStats(C.Structure):
    _fields_ = ([(n, C.c_double) for n in PSTAT_SUMMARIES] +
                [("taup", C.c_double), ("St", C.c_double), ("gr_contact", C.c_double),
                 ("npairs_contact", C.c_int64), ("nbins", C.c_int), ("nshells", C.c_int)])


class TileShapes(C.Structure):
    _fields_ = [("sweep_tx", C.c_int), ("sweep_ty", C.c_int), ("rhs_tx", C.c_int), ("rhs_ty", C.c_int),
                ("sweep2_tx", C.c_int), ("sweep2_ty", C.c_int)]


class FlowStat(C.Structure):
    _fields_ = [(n, C.c_double) for n in FLOW_STAT_FIELDS]


class RunStatus(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("steps_done", C.c_int64), ("submit_s", C.c_double),
                ("start_s", C.c_double), ("finish_s", C.c_double), ("bytes", C.c_size_t)]


RUN_STATUS = {0: "queued", 1: "live", 2: "done", 3: "failed"}

P = C.c_void_p
I = C.c_int
D = C.c_double

# name -> argtypes (all return int unless noted)
SIGNATURES = {
    "lcs_ctx_create": [I, C.POINTER(P)],
    "lcs_ctx_destroy": [P],
    "lcs_box_create": [P, C.POINTER(BoxDesc), C.POINTER(P)],
    "lcs_box_destroy": [P],
    "lcs_box_bytes": [P, C.POINTER(C.c_size_t)],
    "lcs_box_describe": [P, C.POINTER(BoxDesc), C.POINTER(I)],
    "lcs_field_upload": [P, I, P, C.c_size_t],
    "lcs_field_download": [P, I, P, C.c_size_t],
    "lcs_fill_ghost": [P, C.c_uint, I, I],
    "lcs_flow_timestep": [P, D, D, C.POINTER(D), P],
    "lcs_rhs": [P, I],
    "lcs_update": [P, I, D],
    "lcs_rhs_update": [P, I, D],
    "lcs_hsmac": [P, I, D, C.POINTER(HsmacParams), C.POINTER(I), C.POINTER(D)],
    "lcs_force_linear": [P, D, D, C.POINTER(D)],
    "lcs_force_rcf": [P, D, I, I, I, D, C.POINTER(D)],
    "lcs_force_rcf_noalloc": [P, D, I, I, I, D, C.POINTER(D)],
    "lcs_boxmean": [P, I, P],
    "lcs_flow_stat": [P, I, C.POINTER(FlowStat)],
    "lcs_autotune": [P, C.POINTER(TileShapes)],
    "lcs_particle_stat": [P, I, C.POINTER(ParticleStatParams), C.c_double, C.POINTER(ParticleStats), P, P, P, P, P],
    "lcs_particle_density": [P, I, P],
    "lcs_set_tiles": [P, C.POINTER(TileShapes)],
    "lcs_get_tiles": [P, C.POINTER(TileShapes)],
    "lcs_particles_add": [P, C.POINTER(ParticleDesc), C.POINTER(I)],
    "lcs_particles_describe": [P, I, C.POINTER(ParticleDesc)],
    "lcs_particles_upload": [P, I, I, P, C.c_size_t],
    "lcs_particles_download": [P, I, I, P, C.c_size_t],
    "lcs_particles_boundary": [P, I, I],
    "lcs_particles_motion": [P, I, I],
    "lcs_particles_update": [P, I, I, D],
    "lcs_particles_timestep": [P, I, D, C.POINTER(D)],
    "lcs_particles_makeindex": [P, I, I],
    "lcs_particles_localize": [P, I],
    "lcs_particles_cellindex": [P, I, P, P, P, C.POINTER(C.c_int64)],
    "lcs_init_spectrum": [P, C.c_uint64, D, D],
    "lcs_init_particles": [P, I, C.c_uint64, D, D, D, I],
    "lcs_step": [P, C.POINTER(StepConfig), C.POINTER(StepLog)],
    "lcs_step_log_read": [P, C.POINTER(StepLog)],
    "lcs_set_state": [P, C.c_int64, D],
    "lcs_batch_create": [P, C.POINTER(P)],
    "lcs_batch_destroy": [P],
    "lcs_batch_add": [P, P, C.POINTER(StepConfig)],
    "lcs_batch_remove": [P, P],
    "lcs_batch_size": [P, C.POINTER(I)],
    "lcs_batch_step": [P, P, I],
    "lcs_batch_sync": [P],
    "lcs_batch_set_graphs": [P, I],
    "lcs_batch_set_persistent": [P, I],
    "lcs_batch_autotune": [P],
    "lcs_batch_graph_stats": [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)],
    "lcs_batch_set_tiles": [P, I, I, I, I],
    "lcs_batch_get_tiles": [P, I, C.POINTER(I), C.POINTER(I), C.POINTER(I)],
    "lcs_batch_set_rhs_tiles": [P, I, I, I],
    "lcs_batch_get_rhs_tiles": [P, I, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "lcs_server_create": [P, C.POINTER(ServerConfig), C.POINTER(P)],
    "lcs_server_destroy": [P],
    "lcs_server_submit": [P, C.POINTER(Request), C.POINTER(C.c_int64)],
    "lcs_server_step": [P, C.POINTER(I), C.POINTER(I)],
    "lcs_server_status": [P, C.c_int64, C.POINTER(RunStatus)],
    "lcs_server_counts": [P, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_size_t)],
    "lcs_server_box": [P, C.c_int64, C.POINTER(P)],
    "lcs_server_release": [P, C.c_int64],
    "lcs_profile": [P, I],
    "lcs_profile_read": [P, I, C.c_char_p, C.c_size_t, C.POINTER(D), C.POINTER(C.c_int64), C.POINTER(D)],
    "lcs_sync": [P],
    "lcs_last_error": None,
    "lcs_version": None,
}

_lib = None


def load():
    """Load the in-tree shared library (raises if it is absent: no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is missing; build it with `make -C lcs.jl_amd` or __graft_entry__.build()")
    L = C.CDLL(LIBPATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        if args is None:
            f.restype = C.c_char_p
            f.argtypes = []
        else:
            f.restype = C.c_int
            f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != LCS_OK:
        msg = load().lcs_last_error().decode()
        if rc == LCS_ERR_INVALID:
            raise LCSArgumentError(rc, msg)
        raise LCSError(rc, msg)
    return rc
