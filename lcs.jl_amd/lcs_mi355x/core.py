"""Host-side mirror of the reference's buffer and phase API over the C ABI.

``Box`` plays ``LCS.Buffer`` (src/buffer.jl:13-34): it owns one periodic box's
fields (Us, Us2, P) and particle profiles in HBM.  The module-level functions
carry the reference's phase names and argument meaning:

=====================  ==========================================================
``fill_ghost``         ``Topologies.synchalo!`` single process (src/topologies.jl:36-45)
``flow_timestep``      ``Flows.timestep`` (src/flows/timestep.jl:6-65)
``rhs``                ``Flows.rhs!`` (src/flows/rhs.jl:9-81)
``update``             ``Flows.update!`` (src/flows/update.jl:12-54)
``couple``             ``Flows.couple!`` with ``HSMAC`` (src/flows/couples/hsmac.jl)
``forcing``            ``Flows.forcing!`` (src/flows/forcings/{lf,rcf}.jl)
``boundary``/``motion``/``particle_update``/``particle_timestep``/``makeindex``
                       ``Particles.*`` (src/particles/*.jl)
``timestep``/``evolve`` ``LCS.timestep!`` + ``LCS.evolve!`` (src/timestep.jl, src/evolve.jl)
=====================  ==========================================================

Errors raise ``LCSArgumentError`` (a ``ValueError``) for invalid arguments, as
the reference raises ``ArgumentError``, and ``LCSError`` otherwise.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FIELDS, PROP_ID, check
from .config import Config, ConstantPowerLF, EnergyPreserveRCF, HSMAC, NoForcing

_ctx_cache = {}


def context(device=0):
    """One lcs_ctx per device, created on first use."""
    if device not in _ctx_cache:
        L = _lib.load()
        h = C.c_void_p()
        check(L.lcs_ctx_create(device, C.byref(h)))
        _ctx_cache[device] = h
    return _ctx_cache[device]


def _np_dtype(dtype):
    d = np.dtype({"f64": np.float64, "f32": np.float32}.get(dtype, dtype))
    if d not in (np.float64, np.float32):
        raise ValueError("dtype must be float64 or float32")
    return d


class Box:
    """One periodic box resident in HBM (LCS.Buffer for one process)."""

    def __init__(self, dims, halo_size=3, params=None, dtype=np.float64, device=0):
        from .config import FlowParams
        self.L = _lib.load()
        self.dims = tuple(int(n) for n in dims)
        self.h = int(halo_size)
        self.dtype = _np_dtype(dtype)
        self.params = params or FlowParams(1.0, 1.0, 1.0)
        desc = _lib.BoxDesc(*self.dims, self.h, _lib.LCS_F64 if self.dtype == np.float64 else _lib.LCS_F32,
                            self.params.U0_m_s, self.params.L0_m, self.params.NU_m2_s)
        h = C.c_void_p()
        check(self.L.lcs_box_create(context(device), C.byref(desc), C.byref(h)))
        self.handle = h
        self.profiles = []

    # -- memory -------------------------------------------------------------
    @property
    def shape(self):
        return tuple(n + 2 * self.h for n in self.dims)

    def nbytes(self):
        v = C.c_size_t()
        check(self.L.lcs_box_bytes(self.handle, C.byref(v)))
        return v.value

    def upload(self, name, A):
        fid = FIELDS[name]
        shape = self.dims if fid >= FIELDS["dU"] else self.shape
        A = np.asfortranarray(A, dtype=self.dtype)
        if A.shape != shape:
            raise ValueError(f"{name}: expected shape {shape}, got {A.shape}")
        check(self.L.lcs_field_upload(self.handle, fid, A.ctypes.data, A.nbytes))

    def download(self, name):
        fid = FIELDS[name]
        shape = self.dims if fid >= FIELDS["dU"] else self.shape
        A = np.empty(shape, dtype=self.dtype, order="F")
        check(self.L.lcs_field_download(self.handle, fid, A.ctypes.data, A.nbytes))
        return A

    def interior(self, name):
        A = self.download(name)
        h = self.h
        return A[h:-h, h:-h, h:-h] if h else A

    def sync(self):
        check(self.L.lcs_sync(self.handle))

    def profile(self, enable=True):
        """Start (or stop and clear) live per-kernel timing on the box stream."""
        check(self.L.lcs_profile(self.handle, int(bool(enable))))

    def profile_read(self):
        """{kernel: (total_ms, launches, algorithmic_bytes)} since profile(True)."""
        out = {}
        i = 0
        buf = C.create_string_buffer(64)
        while True:
            ms = C.c_double()
            n = C.c_int64()
            by = C.c_double()
            rc = self.L.lcs_profile_read(self.handle, i, buf, 64, C.byref(ms), C.byref(n), C.byref(by))
            if rc != 0:
                break
            out[buf.value.decode()] = (ms.value, n.value, by.value)
            i += 1
        return out

    def close(self):
        if getattr(self, "handle", None):
            self.L.lcs_box_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- particles ---------------------------------------------------------
    def add_particles(self, nvalid, ntotal=None, gravity=0.0, density_ratio=843.0, drag="linear", grid_ratio=4):
        ntotal = nvalid if ntotal is None else ntotal
        d = _lib.ParticleDesc(int(nvalid), int(ntotal), float(gravity), float(density_ratio), _lib.DRAG[drag],
                              int(grid_ratio))
        idx = C.c_int()
        check(self.L.lcs_particles_add(self.handle, C.byref(d), C.byref(idx)))
        self.profiles.append(dict(nvalid=int(nvalid), ntotal=int(ntotal), grid_ratio=int(grid_ratio)))
        return idx.value

    def _pdt(self, prop):
        return np.int64 if prop == "ids" else self.dtype

    def upload_prop(self, profile, prop, a):
        nt = self.profiles[profile]["ntotal"]
        a = np.ascontiguousarray(a, dtype=self._pdt(prop))
        if a.shape != (nt,):
            raise ValueError(f"{prop}: expected shape ({nt},), got {a.shape}")
        check(self.L.lcs_particles_upload(self.handle, profile, PROP_ID[prop], a.ctypes.data, a.nbytes))

    def download_prop(self, profile, prop):
        nt = self.profiles[profile]["ntotal"]
        a = np.empty(nt, dtype=self._pdt(prop))
        check(self.L.lcs_particles_download(self.handle, profile, PROP_ID[prop], a.ctypes.data, a.nbytes))
        return a

    def cellindex(self, profile):
        nv = self.profiles[profile]["nvalid"]
        n = C.c_int64()
        check(self.L.lcs_particles_cellindex(self.handle, profile, None, None, None, C.byref(n)))
        hashes = np.zeros(max(nv, 1), np.int64)
        starts = np.zeros(n.value, np.int64)
        stops = np.zeros(n.value, np.int64)
        check(self.L.lcs_particles_cellindex(self.handle, profile, hashes.ctypes.data, starts.ctypes.data,
                                             stops.ctypes.data, None))
        return hashes[:nv], starts, stops

    # -- initial conditions (builder-defined) ------------------------------------
    def init_spectrum(self, seed, k0=4.0, uprime=1.0):
        check(self.L.lcs_init_spectrum(self.handle, int(seed), float(k0), float(uprime)))

    def init_particles(self, profile, seed, st_min=1.0, st_max=1.0, eps=0.5, zero_slip=True):
        check(self.L.lcs_init_particles(self.handle, profile, int(seed), float(st_min), float(st_max), float(eps),
                                        int(bool(zero_slip))))


# ---------------------------------------------------------------------------
# Phase functions
# ---------------------------------------------------------------------------
def fill_ghost(box, names, width, full=True):
    mask = 0
    for n in names:
        mask |= 1 << FIELDS[n]
    check(box.L.lcs_fill_ghost(box.handle, mask, int(width), int(bool(full))))


def flow_timestep(box, cfl, dtmax):
    dt = C.c_double()
    um = (C.c_double * 3)()
    check(box.L.lcs_flow_timestep(box.handle, float(cfl), float(dtmax), C.byref(dt), um))
    return dt.value, tuple(um)


def rhs(box, stage):
    check(box.L.lcs_rhs(box.handle, int(stage)))


def update(box, stage, dt_stage):
    check(box.L.lcs_update(box.handle, int(stage), float(dt_stage)))


def rhs_update(box, stage, dt_stage):
    check(box.L.lcs_rhs_update(box.handle, int(stage), float(dt_stage)))


def _hsmac_params(c):
    return _lib.HsmacParams(int(c.time_blocking), int(c.overlap), c.coloring_id, float(c.omega), float(c.epsp),
                            int(c.itrmin), int(c.itrmax))


def couple(box, hsmac: HSMAC, stage, dt_stage):
    """Returns the reference's (; itr, div)."""
    p = _hsmac_params(hsmac)
    itr = C.c_int()
    div = C.c_double()
    check(box.L.lcs_hsmac(box.handle, int(stage), float(dt_stage), C.byref(p), C.byref(itr), C.byref(div)))
    return {"itr": itr.value, "div": div.value}


def forcing(box, f, dt):
    p = C.c_double()
    if isinstance(f, ConstantPowerLF):
        check(box.L.lcs_force_linear(box.handle, float(f.power), float(dt), C.byref(p)))
        return {"p": p.value}
    if isinstance(f, EnergyPreserveRCF):
        check(box.L.lcs_force_rcf(box.handle, float(f.energy), int(f.kmin), int(f.kmax), int(f.coarse_size),
                                  float(dt), C.byref(p)))
        return {"p": p.value, "e10": f.energy}
    if isinstance(f, NoForcing):
        return {}
    raise ValueError(f"unknown forcing {f!r}")


def boxmean(box, coarse_size):
    c = int(coarse_size)
    out = np.zeros(3 * c ** 3)
    check(box.L.lcs_boxmean(box.handle, c, out.ctypes.data))
    return [out[a * c ** 3:(a + 1) * c ** 3].reshape((c, c, c), order="F") for a in range(3)]


SPECTRUM_WAVENUMBERS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768)


def particle_stat(box, profile, tau_eta, contact_tolerance=0.01, bin_size=16, min_radius=1.25e-4,
                  search_cell_size=2, max_wavenumber=768, nsamples=64):
    """Particles.stat (src/particles/stat.jl:80-155) of one profile: re-bins it
    (makeindex!), then summaries, taup/St, the radial distribution function
    over the cell index and the number-density spectrum.  Defaults are those of
    cfgs/N1024-1.lcs-yaml."""
    prm = _lib.ParticleStatParams(float(contact_tolerance), int(bin_size), float(min_radius), int(search_cell_size),
                                  int(max_wavenumber), int(nsamples))
    nsh = sum(1 for k in SPECTRUM_WAVENUMBERS if k <= max_wavenumber) if max_wavenumber > 0 else 0
    edges = np.zeros(bin_size + 1)
    npairs = np.zeros(bin_size, np.int64)
    gr = np.zeros(bin_size)
    wn = np.zeros(max(nsh, 1))
    enp = np.zeros(max(nsh, 1))
    st = _lib.ParticleStats()
    check(box.L.lcs_particle_stat(box.handle, int(profile), C.byref(prm), float(tau_eta), C.byref(st),
                                  edges.ctypes.data, npairs.ctypes.data, gr.ctypes.data, wn.ctypes.data,
                                  enp.ctypes.data))
    out = {c: {k: getattr(st, f"{c}_{k}") for k in ("mean", "max", "min")} for c in ("x", "y", "z", "u", "v", "w", "diam")}
    out.update(taup=st.taup, St=st.St, rdf=dict(gr_contact=st.gr_contact, npairs_contact=st.npairs_contact,
                                                 edges=edges, gr=gr, npairs=npairs),
               spectrum=dict(wavenumbers=wn[:nsh], Enp=enp[:nsh]))
    return out


def particle_density(box, profile):
    """Particles.density (src/particles/stat.jl:171-217): re-bins the profile,
    then the SPH density of every valid particle in the re-binned order."""
    nv = box.profiles[profile]["nvalid"]
    out = np.zeros(max(nv, 1))
    check(box.L.lcs_particle_density(box.handle, int(profile), out.ctypes.data))
    return out[:nv]


def _shapes(t):
    return {"sweep": (t.sweep_tx, t.sweep_ty), "rhs": (t.rhs_tx, t.rhs_ty), "sweep2": (t.sweep2_tx, t.sweep2_ty)}


def autotune(box):
    """Time the candidate tile shapes of the fused RHS and sweep kernels for
    this box's size and keep the fastest (cached per size).  No result
    changes.  Returns {"sweep": (tx, ty), "rhs": (tx, ty), "sweep2": (tx, ty)}."""
    t = _lib.TileShapes()
    check(box.L.lcs_autotune(box.handle, C.byref(t)))
    return _shapes(t)


def set_tiles(box, sweep=None, rhs=None, sweep2=None):
    """Force tile shapes (sweep2=(0, 0): no two-sweep pass); with no
    arguments, return to the size rule."""
    if sweep is None and rhs is None and sweep2 is None:
        check(box.L.lcs_set_tiles(box.handle, None))
        return
    cur = get_tiles(box)
    sweep = sweep or cur["sweep"]
    rhs = rhs or cur["rhs"]
    sweep2 = cur["sweep2"] if sweep2 is None else sweep2
    t = _lib.TileShapes(int(sweep[0]), int(sweep[1]), int(rhs[0]), int(rhs[1]), int(sweep2[0]), int(sweep2[1]))
    check(box.L.lcs_set_tiles(box.handle, C.byref(t)))


def get_tiles(box):
    t = _lib.TileShapes()
    check(box.L.lcs_get_tiles(box.handle, C.byref(t)))
    return _shapes(t)


def flow_stat(box, integral_length=False):
    """Flows.stat (src/flows/stat.jl:53-155): statistics of the box's current
    velocity (after a FullHalo ghost fill).  Returns a dict with the
    reference's FlowStat fields; u, v, w are (mean, max, min) summaries."""
    st = _lib.FlowStat()
    check(box.L.lcs_flow_stat(box.handle, int(bool(integral_length)), C.byref(st)))
    d = {n: getattr(st, n) for n in _lib.FLOW_STAT_FIELDS}
    out = {c: {"mean": d.pop(c + "_mean"), "max": d.pop(c + "_max"), "min": d.pop(c + "_min")} for c in "uvw"}
    out.update(d)
    return out


def boundary(box, profile, stage):
    check(box.L.lcs_particles_boundary(box.handle, profile, int(stage)))


def motion(box, profile, stage):
    check(box.L.lcs_particles_motion(box.handle, profile, int(stage)))


def particle_update(box, profile, stage, dt_stage):
    check(box.L.lcs_particles_update(box.handle, profile, int(stage), float(dt_stage)))


def particle_timestep(box, profile, dtmax):
    dt = C.c_double()
    check(box.L.lcs_particles_timestep(box.handle, profile, float(dtmax), C.byref(dt)))
    return dt.value


def makeindex(box, profile, stage=1):
    check(box.L.lcs_particles_makeindex(box.handle, profile, int(stage)))


def localize(box, profile):
    """Reorder the profile's device storage by position for gather locality.
    Changes no result; upload/download keep the logical order."""
    check(box.L.lcs_particles_localize(box.handle, profile))


# ---------------------------------------------------------------------------
# Whole step
# ---------------------------------------------------------------------------
def step_config(config: Config, localize_interval=0):
    """The C step config of a Config; localize_interval > 0 reorders particle
    storage by position every that many steps (no effect on results)."""
    f = config.flow.forcing
    fk = _lib.Forcing(_lib.FORCING[f.kind], getattr(f, "power", 0.0), getattr(f, "energy", 0.0),
                      getattr(f, "kmin", 0), getattr(f, "kmax", 0), getattr(f, "coarse_size", 0))
    return _lib.StepConfig(config.timestep.cfl, config.timestep.dtmax, _hsmac_params(config.flow.couple), fk,
                           int(localize_interval))


def step(box, scfg, want_log=True):
    """LCS.timestep! + LCS.evolve! on the device; returns the step log."""
    log = _lib.StepLog()
    check(box.L.lcs_step(box.handle, C.byref(scfg), C.byref(log) if want_log else None))
    if not want_log:
        return None
    return log_dict(log)


class Batch:
    """An ensemble of independent boxes advanced together, one kernel launch
    per phase for all of them (lcs_batch_*).  No counterpart in the reference,
    whose ensembles are separate runs; each box advances as ``step`` would
    advance it alone, and a box's result does not depend on the other boxes."""

    def __init__(self, device=0):
        self.L = _lib.load()
        h = C.c_void_p()
        check(self.L.lcs_batch_create(context(device), C.byref(h)))
        self.handle = h
        self.boxes = []
        self._cfgs = []

    def add(self, box, scfg):
        check(self.L.lcs_batch_add(self.handle, box.handle, C.byref(scfg)))
        self.boxes.append(box)
        self._cfgs.append(scfg)

    def remove(self, box):
        check(self.L.lcs_batch_remove(self.handle, box.handle))
        i = self.boxes.index(box)
        del self.boxes[i], self._cfgs[i]

    def __len__(self):
        n = C.c_int()
        check(self.L.lcs_batch_size(self.handle, C.byref(n)))
        return n.value

    def step(self, want_log=True):
        """One timestep! + evolve! of every box; returns one log per box."""
        n = len(self)
        if not want_log:
            check(self.L.lcs_batch_step(self.handle, None, 0))
            return None
        logs = (_lib.StepLog * max(n, 1))()
        check(self.L.lcs_batch_step(self.handle, logs, n))
        return [log_dict(logs[i]) for i in range(n)]

    def set_graphs(self, on):
        """hipGraph mode of the batched step (lcs_batch_set_graphs): True or 1
        one graph per step once warm (default), 2 the runs between polls only,
        3 one graph per step with the fewest HSMAC groups, False or 0 eager."""
        check(self.L.lcs_batch_set_graphs(self.handle, int(on)))

    def set_persistent(self, on, cus=None):
        """Persistent sweep grids over the tile table (lcs_batch_set_persistent):
        True walks the tiles with as many workgroups as the device holds at
        once (or as `cus` CUs hold), False (default) launches one workgroup
        per tile."""
        n = 0 if not on else (-1 if cus is None else int(cus))
        check(self.L.lcs_batch_set_persistent(self.handle, n))

    def graph_stats(self):
        """{"whole", "gated", "segmented"}: steps by path (lcs_batch_graph_stats)."""
        v = [C.c_int64() for _ in range(3)]
        check(self.L.lcs_batch_graph_stats(self.handle, *[C.byref(x) for x in v]))
        return dict(zip(("whole", "gated", "segmented"), (x.value for x in v)))

    def autotune(self):
        """Picks the batched sweep's tile shape and z-chunk, and the fused RHS's
        tile shape, per size class by timing (lcs_batch_autotune); no box
        changes.  Returns tiles()."""
        check(self.L.lcs_batch_autotune(self.handle))
        return self.tiles()

    def set_tiles(self, i, tx, ty, zchunk):
        """Tiling of the size class of box i: tile tx x ty, z-chunks of about zchunk planes."""
        check(self.L.lcs_batch_set_tiles(self.handle, int(i), int(tx), int(ty), int(zchunk)))

    def tiles(self):
        """[(tx, ty, z-chunk planes)] of every box's sweep tiling, in order of addition."""
        out = []
        for i in range(len(self)):
            v = [C.c_int() for _ in range(3)]
            check(self.L.lcs_batch_get_tiles(self.handle, i, *[C.byref(x) for x in v]))
            out.append(tuple(x.value for x in v))
        return out

    def set_rhs_tiles(self, i, tx, ty):
        """Fused RHS tile of the size class of box i (lcs_batch_set_rhs_tiles)."""
        check(self.L.lcs_batch_set_rhs_tiles(self.handle, int(i), int(tx), int(ty)))

    def rhs_tiles(self):
        """[(tx, ty)] of every box's fused RHS tile, in order of addition."""
        out = []
        for i in range(len(self)):
            v = [C.c_int() for _ in range(2)]
            check(self.L.lcs_batch_get_rhs_tiles(self.handle, i, *[C.byref(x) for x in v]))
            out.append(tuple(x.value for x in v))
        return out

    def sync(self):
        check(self.L.lcs_batch_sync(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            self.L.lcs_batch_destroy(self.handle)
            self.handle = None
            self.boxes, self._cfgs = [], []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BorrowedBox(Box):
    """A box owned by the server (valid until the request is released)."""

    def __init__(self, handle, dims, dtype, particles=0):
        self.L = _lib.load()
        self.handle = handle
        self.dims = tuple(dims)
        self.h = 3
        self.dtype = _np_dtype(dtype)
        # a served run holds one particle profile (grid ratio 4) when it has particles
        self.profiles = [dict(nvalid=int(particles), ntotal=int(particles), grid_ratio=4)] if particles > 0 else []

    def close(self):
        self.handle = None


class Server:
    """Serving loop over one batch (lcs_server_*): runs are submitted as
    requests, admitted by slot and device-memory footprint, advanced together
    one batched step per ``step()``, and retired when they reach their step
    count.  No counterpart in the reference.  memory_budget 0 takes nine
    tenths of the device's free memory (at most 288e9 bytes); a request larger
    than the budget is refused with LCSError (capacity)."""

    def __init__(self, memory_budget=0, max_live=128, cfl=0.3, dtmax=0.01, hsmac=None,
                 localize_interval=10, device=0):
        self.L = _lib.load()
        hs = hsmac or HSMAC(True, True, "red-black-fast", 1.7, 1e-3, 6, 300)
        cfg = _lib.ServerConfig(int(memory_budget), int(max_live), float(cfl), float(dtmax), _hsmac_params(hs),
                                int(localize_interval))
        h = C.c_void_p()
        check(self.L.lcs_server_create(context(device), C.byref(cfg), C.byref(h)))
        self.handle = h
        self._req = {}

    def submit(self, n, steps, seed, particles=0, st=(1.0, 1.0), forcing="linear-power", forcing_param=0.5,
               k0=4.0, Re=53.3, dtype="f32"):
        r = _lib.Request(int(n), _lib.LCS_F64 if _np_dtype(dtype) == np.float64 else _lib.LCS_F32, int(particles),
                         float(st[0]), float(st[1]), _lib.FORCING[forcing], float(forcing_param), int(steps),
                         int(seed) & (2**64 - 1), float(k0), float(Re))
        i = C.c_int64()
        check(self.L.lcs_server_submit(self.handle, C.byref(r), C.byref(i)))
        self._req[i.value] = (int(n), _np_dtype(dtype), int(particles))
        return i.value

    def step(self):
        """Admit what fits, advance every live run one step; returns (live, finished)."""
        live, fin = C.c_int(), C.c_int()
        check(self.L.lcs_server_step(self.handle, C.byref(live), C.byref(fin)))
        return live.value, fin.value

    def status(self, rid):
        s = _lib.RunStatus()
        check(self.L.lcs_server_status(self.handle, int(rid), C.byref(s)))
        return {"id": s.id, "status": _lib.RUN_STATUS[s.status], "steps_done": s.steps_done,
                "submit_s": s.submit_s, "start_s": s.start_s, "finish_s": s.finish_s, "bytes": s.bytes}

    def counts(self):
        q, lv, m = C.c_int(), C.c_int(), C.c_size_t()
        check(self.L.lcs_server_counts(self.handle, C.byref(q), C.byref(lv), C.byref(m)))
        return {"queued": q.value, "live": lv.value, "memory_used": m.value}

    def box(self, rid):
        """The run's box (fields and particles readable until release)."""
        h = C.c_void_p()
        check(self.L.lcs_server_box(self.handle, int(rid), C.byref(h)))
        n, dt, npart = self._req[rid]
        return _BorrowedBox(h, (n, n, n), dt, npart)

    def release(self, rid):
        check(self.L.lcs_server_release(self.handle, int(rid)))
        self._req.pop(rid, None)

    def close(self):
        if getattr(self, "handle", None):
            self.L.lcs_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_state(box, step, t):
    """state.step = step; state.t = t (LCSIO.load_resume_state!,
    src/LCSIO/load.jl:1-21)."""
    check(box.L.lcs_set_state(box.handle, int(step), float(t)))


def read_log(box):
    log = _lib.StepLog()
    check(box.L.lcs_step_log_read(box.handle, C.byref(log)))
    return log_dict(log)


def log_dict(log):
    return {"step": log.step, "t": log.t, "dt": log.dt,
            "flow": {"couple1": {"itr": log.itr1, "div": log.div1},
                     "forcing": {"p": log.p, "e10": log.e10},
                     "couple2": {"itr": log.itr2, "div": log.div2}}}
