"""LCSIO for the mi355x backend: output directory and resume policy, restart
and statistics files in the reference's HDF5 layout (src/LCSIO/*.jl).

The files are written and read by liblcs_io.so (include/lcs_io.h) through the
HDF5 C library; this module is the host-side mirror of the reference's
functions with the same names.  A restart file written here loads in the
reference (HDF5.jl) and the other way round:

* ``restart.h5.<step>``: ``flow/{U,V,W,P}`` (global nx*ny*nz, no halo) and
  ``particles/<i>/{counts,id,x,y,z,u,v,w,diam}`` (src/LCSIO/save.jl:40-167);
* ``stat.h5``: one group per step with ``state``, ``stat`` and ``log``
  written by ``deepwrite`` (save.jl:14-33, src/utils/hdf5.jl:5-19).
"""
import ctypes as C
import os
import re

import numpy as np
import yaml

from . import _lib
from .config import Config, to_dict

FILE_CONFIG = "config.lcs-yaml"  # src/LCSIO/LCSIO.jl:17-20
FILE_STAT = "stat.h5"
FILE_RESTART_REGEX = re.compile(r"restart\.h5\.(\d+)$")
FILE_RESTART_FMT = "restart.h5.{:d}"

TAG_STATE, TAG_STAT, TAG_LOG = "state", "stat", "log"
TAG_FLOW = "flow"
TAG_PARTICLE = "particles"

ENV_LCS_RESUME = "LCS_RESUME"  # LCS.ENV_LCS_RESUME

H5_I64, H5_F64, H5_F32, H5_STR = 0, 1, 2, 3
H5_NONE, H5_GROUP, H5_DATASET = 0, 1, 2

IOLIBPATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "liblcs_io.so")


class H5Entry(C.Structure):
    _fields_ = [("path", C.c_char_p), ("type", C.c_int), ("rank", C.c_int), ("dims", C.c_int64 * 3),
                ("data", C.c_void_p)]


P = C.c_void_p
I = C.c_int
IO_SIGNATURES = {
    "lcs_save_restart": [P, C.c_char_p],
    "lcs_load_flow_fields": [P, C.c_char_p],
    "lcs_load_particles": [P, C.c_char_p, I],
    "lcs_save_stat": [C.c_char_p, C.c_int64, C.POINTER(H5Entry), I],
    "lcs_h5_write": [C.c_char_p, C.POINTER(H5Entry), I],
    "lcs_h5_info": [C.c_char_p, C.c_char_p, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(C.c_int64)],
    "lcs_h5_read": [C.c_char_p, C.c_char_p, I, P, C.c_int64],
    "lcs_h5_keys": [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(I), C.POINTER(C.c_size_t)],
    "lcs_io_last_error": None,
}

_io = None


def load_io():
    """Load the in-tree liblcs_io.so (raises if it is absent: it is built only
    where the HDF5 headers are found)."""
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IOLIBPATH):
        raise ImportError(f"{IOLIBPATH} is missing; it is built by `make -C lcs.jl_amd` where HDF5 is installed")
    _lib.load()  # the IO library resolves the compute library's symbols
    L = C.CDLL(IOLIBPATH)
    for name, args in IO_SIGNATURES.items():
        f = getattr(L, name)
        if args is None:
            f.restype = C.c_char_p
            f.argtypes = []
        else:
            f.restype = C.c_int
            f.argtypes = args
    _io = L
    return L


def _check(rc):
    if rc != _lib.LCS_OK:
        msg = load_io().lcs_io_last_error().decode()
        if rc == _lib.LCS_ERR_INVALID:
            raise _lib.LCSArgumentError(rc, msg)
        raise _lib.LCSError(rc, msg)


def _b(s):
    return os.fspath(s).encode()


# ---- deepwrite / deepread ---------------------------------------------------
def _entries(tree, prefix=""):
    """Leaves of Utils.deepwrite (src/utils/hdf5.jl:5-19) with PolySerde.normalize
    (lib/polyserde/src/*.jl:185-200) applied: mappings recurse by key, lists and
    tuples of non-numbers by 1-based index, numeric arrays are datasets, None is
    skipped.  Returns (path, type, rank, dims, ndarray|bytes) tuples."""
    out = []
    if tree is None:
        return out
    if isinstance(tree, dict):
        for k, v in tree.items():
            out += _entries(v, f"{prefix}{k}/")
        return out
    path = prefix.rstrip("/")
    if isinstance(tree, str):
        return [(path, H5_STR, 0, (len(tree.encode()), 0, 0), tree.encode())]
    if isinstance(tree, (bool, np.bool_)):
        raise TypeError(f"cannot write value of type bool at path {path}")
    if isinstance(tree, (int, np.integer)):
        return [(path, H5_I64, 0, (0, 0, 0), np.array(int(tree), np.int64))]
    if isinstance(tree, (float, np.floating)):
        t = H5_F32 if isinstance(tree, np.float32) else H5_F64
        return [(path, t, 0, (0, 0, 0), np.array(tree, np.float32 if t == H5_F32 else np.float64))]
    if isinstance(tree, (list, tuple)) and not all(isinstance(v, (int, float, np.number)) and
                                                   not isinstance(v, bool) for v in tree):
        for i, v in enumerate(tree):
            out += _entries(v, f"{prefix}{i + 1}/")
        return out
    a = np.asarray(tree)
    if a.dtype.kind in "iu":
        a, t = np.ascontiguousarray(a, np.int64), H5_I64
    elif a.dtype == np.float32:
        a, t = np.ascontiguousarray(a), H5_F32
    elif a.dtype.kind == "f":
        a, t = np.ascontiguousarray(a, np.float64), H5_F64
    else:
        raise TypeError(f"cannot write value of type {a.dtype} at path {path}")
    if a.ndim > 3:
        raise TypeError(f"cannot write a {a.ndim}-d array at path {path}")
    # a Julia array of size (n1, n2, n3) is the C array [n3][n2][n1]
    dims = tuple(reversed(a.shape)) + (0,) * (3 - a.ndim)
    return [(path, t, a.ndim, dims, a)]


def _entry_array(items):
    arr = (H5Entry * max(len(items), 1))()
    keep = []
    for q, (path, t, rank, dims, data) in enumerate(items):
        pb = path.encode()
        keep += [pb, data]
        ptr = C.cast(C.c_char_p(data), C.c_void_p) if isinstance(data, bytes) else data.ctypes.data
        arr[q] = H5Entry(pb, t, rank, (C.c_int64 * 3)(*dims), ptr)
    return arr, keep


def h5_write(file, tree):
    """Write a nested mapping of scalars and arrays into file (created when
    absent); an existing dataset is an error.  Arrays are given in C order:
    a numpy array of shape (nz, ny, nx) is stored as the reference's
    (nx, ny, nz) array."""
    items = _entries(tree)
    arr, keep = _entry_array(items)
    _check(load_io().lcs_h5_write(_b(file), arr, len(items)))


def h5_info(file, path):
    """(kind, type, dims) of h[path]; dims in the reference's (Julia) order."""
    kind, typ, rank = I(), I(-1), I(0)
    dims = (C.c_int64 * 8)()
    _check(load_io().lcs_h5_info(_b(file), path.encode(), C.byref(kind), C.byref(typ), C.byref(rank), dims))
    return kind.value, typ.value, tuple(dims[q] for q in range(rank.value))


def haskey(file, path):
    return h5_info(file, path)[0] != H5_NONE


def keys(file, path="/"):
    """keys(h[path]), sorted by name."""
    L = load_io()
    n, need = I(), C.c_size_t()
    rc = L.lcs_h5_keys(_b(file), path.encode(), None, 0, C.byref(n), C.byref(need))
    if rc not in (_lib.LCS_OK, _lib.LCS_ERR_CAPACITY):
        _check(rc)
    buf = C.create_string_buffer(max(need.value, 1))
    _check(L.lcs_h5_keys(_b(file), path.encode(), buf, len(buf), C.byref(n), C.byref(need)))
    return [k.decode() for k in buf.raw[:need.value].split(b"\0")[:n.value]]


def h5_read(file, path):
    """read(h[path]): a Python scalar for a scalar dataset, a numpy array in C
    order (the reference's (nx, ny, nz) array as shape (nz, ny, nx)) otherwise,
    or a str."""
    kind, typ, dims = h5_info(file, path)
    if kind != H5_DATASET:
        raise KeyError(f"{file} has no dataset {path}")
    L = load_io()
    if typ == H5_STR:
        buf = C.create_string_buffer(1 << 16)
        _check(L.lcs_h5_read(_b(file), path.encode(), H5_STR, buf, len(buf)))
        return buf.value.decode()
    dt = {H5_I64: np.int64, H5_F32: np.float32, H5_F64: np.float64}.get(typ)
    if dt is None:
        raise TypeError(f"{path}: unsupported element type")
    shape = tuple(reversed(dims))
    a = np.zeros(max(int(np.prod(shape)), 1), dt)
    _check(L.lcs_h5_read(_b(file), path.encode(), typ, a.ctypes.data, a.size))
    if not shape:
        return a[0].item()
    return a[:int(np.prod(shape))].reshape(shape)


def deepread(file, path="/"):
    """Utils.deepread (src/utils/hdf5.jl:21-30): groups become dicts."""
    kind = h5_info(file, path)[0]
    if kind == H5_GROUP:
        base = path.rstrip("/")
        return {k: deepread(file, f"{base}/{k}" if base else k) for k in keys(file, path)}
    return h5_read(file, path)


# ---- state / stat / log as the reference normalises them -------------------
_FLOW_STAT_NAMES = {"eps": "ϵ", "tau_eta": "τη", "l_mu": "lμ", "l_eta": "lη",
                    "Re_lambda": "Reλ", "k_eta": "kη", "C_eps": "Cϵ"}
_FLOW_STAT_ORDER = ("u", "v", "w", "k", "eps", "tau_eta", "l_mu", "l_eta", "Re_lambda", "urms", "k_eta", "skew",
                    "flat", "L", "C_eps")  # FlowStat fields (src/flows/stat.jl:23-41)


def state_tree(box, log=None):
    """PolySerde.normalize(state) (src/state.jl:15-23): step, t, dt and one
    {nvalid} per particle profile; FlowState has no fields."""
    if log is None:
        from .core import read_log
        log = read_log(box)
    out = {"step": int(log["step"]), "t": float(log["t"]), "dt": float(log["dt"])}
    if box.profiles:
        out["particles"] = [{"nvalid": int(p["nvalid"])} for p in box.profiles]
    return out


def stat_tree(state, flow, particles=()):
    """PolySerde.normalize(Stat(StateStat(state), flow, particles))
    (src/stat.jl:75-105): the reference's field names, Unicode included."""
    f = {(_FLOW_STAT_NAMES.get(k, k)): flow[k] for k in _FLOW_STAT_ORDER}
    ps = []
    for p in particles:
        q = {c: p[c] for c in ("x", "y", "z", "u", "v", "w", "diam")}
        q["τp"] = p["taup"]
        q["St"] = p["St"]
        q["rdf"] = {k: p["rdf"][k] for k in ("gr_contact", "npairs_contact", "edges", "gr", "npairs")}
        q["spectrum"] = {k: p["spectrum"][k] for k in ("wavenumbers", "Enp")}
        ps.append(q)
    return {"state": {"step": state["step"], "t": state["t"], "dt": state["dt"]}, "flow": f, "particles": ps}


def log_tree(log, forcing="linear-power", nvalids=()):
    """PolySerde.normalize of evolve!'s log (src/evolve.jl:127-145).  Phases
    that log nothing (rhs!, update!, motion!, particle update!, no forcing)
    normalise to empty mappings and write nothing.  couple! logs (itr, div)
    (src/flows/couples/hsmac.jl:128), linear forcing (p,) (lf.jl:69), RC
    forcing (p, e10) (rcf.jl:99), and boundary! per profile the
    ExchangeSummary (mean, min, max) of nvalid, nrecv and nsend, which on one
    process are (nvalid, nvalid, nvalid) and zeros (boundary.jl:1-20,54-58)."""
    fl = log["flow"]
    flow = {"couple1": {"itr": int(fl["couple1"]["itr"]), "div": float(fl["couple1"]["div"])}}
    if forcing == "linear-power":
        flow["forcing"] = {"p": float(fl["forcing"]["p"])}
    elif forcing == "rc-energy":
        flow["forcing"] = {"p": float(fl["forcing"]["p"]), "e10": float(fl["forcing"]["e10"])}
    flow["couple2"] = {"itr": int(fl["couple2"]["itr"]), "div": float(fl["couple2"]["div"])}
    out = {}
    if len(nvalids):
        def summary(x):
            return {"mean": int(x), "min": int(x), "max": int(x)}

        b = [{"nvalid": summary(n), "nrecvs": summary(0), "nsends": summary(0)} for n in nvalids]
        out["particles"] = {"boundary1": b, "boundary2": b}
    out["flow"] = flow
    return out


# ---- files --------------------------------------------------------------
def savestat(config_or_file, state, stat, log=None):
    """LCSIO.savestat (src/LCSIO/save.jl:1-33): replaces group "<step>" of
    stat.h5 (in config.outdir, or the file given) with state, stat and log."""
    file = (os.path.join(config_or_file.outdir, FILE_STAT) if isinstance(config_or_file, Config)
            else config_or_file)
    items = _entries({TAG_STATE: state, TAG_STAT: stat, TAG_LOG: log or {}})
    arr, keep = _entry_array(items)
    _check(load_io().lcs_save_stat(_b(file), int(state["step"]), arr, len(items)))


def saverestart(box, config_or_file, step=None):
    """LCSIO.saverestart: with a Config, writes config.outdir/restart.h5.<step>
    (save.jl:35-43; step defaults to the box's state.step); with a path,
    writes that file (save.jl:45-56).  Returns the file name."""
    if isinstance(config_or_file, Config):
        if step is None:
            from .core import read_log
            step = read_log(box)["step"]
        file = os.path.join(config_or_file.outdir, FILE_RESTART_FMT.format(int(step)))
    else:
        file = os.fspath(config_or_file)
    _check(load_io().lcs_save_restart(box.handle, _b(file)))
    return file


def load_flow_fields(file, box):
    """LCSIO.load_flow_fields! (src/LCSIO/load.jl:23-42)."""
    _check(load_io().lcs_load_flow_fields(box.handle, _b(file)))


def load_particles(file, box, profile):
    """LCSIO.load_particle_counts! then load_particle_{id,position,velocity,size}!
    (load.jl:44-118) for profile (0-based; the file's group is profile+1)."""
    _check(load_io().lcs_load_particles(box.handle, _b(file), int(profile)))


# ---- output directory and resume policy (src/LCSIO/config.jl, utils.jl) -----
def find_restart_files(dirname):
    return [os.path.join(dirname, f) for f in sorted(os.listdir(dirname)) if FILE_RESTART_REGEX.search(f)]


def extract_step(restart_file):
    m = FILE_RESTART_REGEX.search(os.fspath(restart_file))
    if m is None:
        raise ValueError(f"invalid restart file name {restart_file}")
    return m.group(1)


def find_latest_restart_file(dirname):
    files = find_restart_files(dirname)
    if not files:
        raise ValueError(f"no restart files found in directory {dirname}")
    return max(files, key=lambda f: int(extract_step(f)))


def isresumable(config):
    if not os.path.isdir(config.outdir):
        return False
    return bool(find_restart_files(config.outdir)) and os.path.isfile(os.path.join(config.outdir, FILE_CONFIG))


def _resume_env():
    return os.environ.get(ENV_LCS_RESUME, "").strip().lower() in ("1", "true", "yes", "y", "t", "on")


def patch_resume(config):
    """patch_resume (src/LCSIO/config.jl:79-110): flow and particles restart
    from the latest restart file."""
    import dataclasses

    from .config import ParticleInit, RestartInit
    f = find_latest_restart_file(config.outdir)
    flow = dataclasses.replace(config.flow, init=RestartInit(f))
    parts = tuple(dataclasses.replace(p, init=ParticleInit(velocity="restart", size="restart", restart_file=f,
                                                           restart_profile=i + 1))
                  for i, p in enumerate(config.particles))
    return dataclasses.replace(config, flow=flow, particles=parts, resume=True)


def init(config):
    """LCSIO.init (src/LCSIO/config.jl:44-77): an existing outdir is an error
    unless LCS_RESUME is set and it is resumable; a new run creates outdir and
    saves its config there."""
    resume = _resume_env()
    if os.path.isdir(config.outdir) and not resume:
        raise ValueError(f"output directory {config.outdir} already exists, set LCS_RESUME=1 to resume")
    if os.path.isdir(config.outdir) and resume:
        if not isresumable(config):
            raise ValueError(f"output directory {config.outdir} exists but is not resumable, missing restart "
                             "files or config.lcs-yaml")
        return patch_resume(config)
    os.makedirs(config.outdir, exist_ok=True)
    with open(os.path.join(config.outdir, FILE_CONFIG), "w") as fh:
        yaml.safe_dump(to_dict(config), fh, sort_keys=False)
    return config


def oninterval(step, interval):
    """LCSIO.oninterval (src/LCSIO/utils.jl:7-11) as a predicate; an interval
    of 0 never fires."""
    return interval > 0 and step % interval == 0


def cleanrestart(step, config):
    """LCSIO.cleanrestart (src/LCSIO/utils.jl:13-22): removes the restart
    file num_restart_files intervals back."""
    if not os.path.isdir(config.outdir):
        raise ValueError(f"output directory {config.outdir} does not exist")
    s = config.simulate
    old = os.path.join(config.outdir, FILE_RESTART_FMT.format(step - s.interval_restart * s.num_restart_files))
    if os.path.isfile(old):
        os.remove(old)


def load_resume_state(config):
    """LCSIO.load_resume_state! (src/LCSIO/load.jl:1-21): (step, t) from the
    latest restart file's name and stat.h5's <step>/state/t."""
    f = find_latest_restart_file(config.outdir)
    step = extract_step(f)
    t = h5_read(os.path.join(config.outdir, FILE_STAT), f"{step}/{TAG_STATE}/t")
    return int(step), float(t)
