"""
    LCSMI355X

The MI355X (gfx950) backend of LCS.jl: `backend: {kind: mi355x}` in a config
selects it, and `LCS.simulate(config)` runs unchanged.  Device state lives in
HBM behind the C ABI of `liblcs_mi355x.so` (header `include/lcs_mi355x.h`);
this package registers the backend kind and specialises the reference's
positional entry points on `LCS.Config{<:Any,MI355XBackend}`:

| reference entry point                              | file:line                          | C call                          |
|:---------------------------------------------------|:-----------------------------------|:--------------------------------|
| `LCS.Buffer(mode, config, topo)`                   | `src/buffer.jl:27-34`              | `lcs_box_create`, `lcs_particles_add` |
| `LCS.init!(buf, state, config, topo)`              | `src/init.jl:6-9`                  | host init, `lcs_field_upload`, `lcs_particles_upload` |
| `Topologies.synchalo!(style, As, buf, grid, backend, topo)` | `src/topologies.jl:36-45` | `lcs_fill_ghost`                |
| `Flows.timestep(buf, state, config, topo)`         | `src/flows/timestep.jl:6`          | `lcs_flow_timestep`             |
| `Particles.timestep(buf, state, config, topo)`     | `src/particles/timestep.jl:6`      | `lcs_particles_timestep`        |
| `Flows.rhs!(buf, state, config, topo, stage)`      | `src/flows/rhs.jl:9`               | `lcs_rhs`                       |
| `Flows.update!(buf, state, config, topo, stage)`   | `src/flows/update.jl:12`           | `lcs_update`                    |
| `Flows.couple!(couple, buf, state, config, topo, stage)` | `src/flows/couples/couple.jl:13-17` | `lcs_hsmac`             |
| `Flows.forcing!(::ConstantPowerLF, ...)`           | `src/flows/forcings/lf.jl:6-13`    | `lcs_force_linear`              |
| `Flows.forcing!(::EnergyPreserveRCF, ...)`         | `src/flows/forcings/rcf.jl:1-8`    | `lcs_force_rcf`                 |
| `Particles.boundary!/motion!/update!/makeindex!`   | `src/particles/boundary.jl:33`, `motion.jl:15`, `update.jl:18`, `makeindex.jl:9` | `lcs_particles_*` |
| `LCS.evolve!(::FlowMode / ::FlowParticleMode, ...)` | `src/evolve.jl:15,63`             | `lcs_step` (one call per step)  |
| `Flows.stat`, `Particles.stat`                     | `src/flows/stat.jl:54`, `src/particles/stat.jl:87` | `lcs_flow_stat`, `lcs_particle_stat` |
| `LCSIO.saverestart(file, ...)`                     | `src/LCSIO/save.jl:45-56`          | `lcs_save_restart` (`liblcs_io.so`) |
| `KA.synchronize(::MI355XBackend)`                  | `src/log.jl:12,33,36`              | `lcs_sync`                      |

`evolve!` makes one `lcs_step` call and rebuilds the reference's log
NamedTuple (`src/evolve.jl:127-145`); the phase overrides serve callers that
step phase by phase (the reference's unit tests, `cb_between_rk_stages`
users).  Initial conditions run the reference's own `init!` on host arrays
(every init kind, restart included) and upload the result.
"""
module LCSMI355X

using Accessors
using KernelAbstractions
using LCS
using Libdl
using Parallel
using PolySerde
using Topologies

const KA = KernelAbstractions
const Flows = LCS.Flows
const Particles = LCS.Particles
const LCSIO = LCS.LCSIO

# the compute and restart libraries, built in-tree by `make -C lcs.jl_amd`
const libmi355x = get(ENV, "LCS_MI355X_LIB",
                      normpath(joinpath(@__DIR__, "..", "..", "..", "lcs.jl_amd", "lcs_mi355x", "liblcs_mi355x.so")))
const libio = get(ENV, "LCS_MI355X_IO_LIB", joinpath(dirname(libmi355x), "liblcs_io.so"))

# ---------------------------------------------------------------------------
# Backend type and YAML kind (as ext/AMDGPUExt.jl:8-14 registers "amdgpu")
# ---------------------------------------------------------------------------
"""
    MI355XBackend <: KA.GPU

Selected by `backend: {kind: mi355x}`.  The device is `LCS_MI355X_DEVICE`
(default 0): one process drives one GPU.
"""
struct MI355XBackend <: KA.GPU end
PolySerde.@kind KA.Backend "mi355x" MI355XBackend
PolySerde.deserialize(::Type{MI355XBackend}, ::PolySerde.SerdeDict) = MI355XBackend()
PolySerde.serialize(::Type{MI355XBackend}, ::MI355XBackend, dicttype::PolySerde.SerdeDictType) = dicttype()

const MI355XConfig = LCS.Config{<:Any,MI355XBackend}

# ---------------------------------------------------------------------------
# C ABI: error classes, structs (field order and types of include/lcs_mi355x.h)
# ---------------------------------------------------------------------------
const LCS_OK = Cint(0)
const LCS_ERR_INVALID = Cint(-1)
const LCS_F64 = Cint(0)
const LCS_F32 = Cint(1)
const FIELD_U, FIELD_V, FIELD_W, FIELD_U2, FIELD_V2, FIELD_W2, FIELD_P = Cint.(0:6)
const FORCING_NONE, FORCING_LINEAR_POWER, FORCING_RC_ENERGY = Cint.(0:2)

lasterror() = unsafe_string(ccall((:lcs_last_error, libmi355x), Cstring, ()))

"""
    check(rc)

Turns a C return code into the reference's exception style: invalid
arguments throw `ArgumentError` (as `boundary.jl:182-187` and `halo.jl:287-288`
do), every other failure an `ErrorException` with the library's message.
"""
function check(rc::Integer)
    rc == LCS_OK && return nothing
    msg = lasterror()
    rc == LCS_ERR_INVALID && throw(ArgumentError(msg))
    throw(ErrorException("liblcs_mi355x ($rc): $msg"))
end

struct BoxDesc
    nx::Cint
    ny::Cint
    nz::Cint
    halo::Cint
    dtype::Cint
    U0_m_s::Cdouble
    L0_m::Cdouble
    NU_m2_s::Cdouble
end

struct HsmacParams
    time_blocking::Cint
    overlap::Cint
    coloring::Cint
    omega::Cdouble
    epsp::Cdouble
    itrmin::Cint
    itrmax::Cint
end

struct ParticleDesc
    nvalid::Int64
    ntotal::Int64
    gravity::Cdouble
    density_ratio::Cdouble
    drag::Cint
    grid_ratio::Cint
end

struct ForcingDesc
    kind::Cint
    power::Cdouble
    energy::Cdouble
    kmin::Cint
    kmax::Cint
    coarse_size::Cint
end

struct StepConfig
    cfl::Cdouble
    dtmax::Cdouble
    hsmac::HsmacParams
    forcing::ForcingDesc
    localize_interval::Cint
end

struct StepLog
    step::Int64
    t::Cdouble
    dt::Cdouble
    itr1::Cint
    itr2::Cint
    div1::Cdouble
    div2::Cdouble
    p::Cdouble
    e10::Cdouble
end
StepLog() = StepLog(0, 0.0, 0.0, 0, 0, 0.0, 0.0, 0.0, 0.0)

struct FlowStats
    u_mean::Cdouble; u_max::Cdouble; u_min::Cdouble
    v_mean::Cdouble; v_max::Cdouble; v_min::Cdouble
    w_mean::Cdouble; w_max::Cdouble; w_min::Cdouble
    k::Cdouble; eps::Cdouble; tau_eta::Cdouble; l_mu::Cdouble; l_eta::Cdouble; Re_lambda::Cdouble
    urms::Cdouble; k_eta::Cdouble; skew::Cdouble; flat::Cdouble; L::Cdouble; C_eps::Cdouble
end

struct ParticleStatParams
    contact_tolerance::Cdouble
    bin_size::Cint
    min_radius::Cdouble
    search_cell_size::Cint
    max_wavenumber::Cint
    nsamples::Cint
end

struct ParticleStats
    x_mean::Cdouble; x_max::Cdouble; x_min::Cdouble
    y_mean::Cdouble; y_max::Cdouble; y_min::Cdouble
    z_mean::Cdouble; z_max::Cdouble; z_min::Cdouble
    u_mean::Cdouble; u_max::Cdouble; u_min::Cdouble
    v_mean::Cdouble; v_max::Cdouble; v_min::Cdouble
    w_mean::Cdouble; w_max::Cdouble; w_min::Cdouble
    diam_mean::Cdouble; diam_max::Cdouble; diam_min::Cdouble
    taup::Cdouble
    St::Cdouble
    gr_contact::Cdouble
    npairs_contact::Int64
    nbins::Cint
    nshells::Cint
end

for T in (BoxDesc, HsmacParams, ParticleDesc, ForcingDesc, StepConfig, StepLog, FlowStats, ParticleStatParams,
          ParticleStats)
    @assert isbitstype(T)
end

# ---------------------------------------------------------------------------
# Device state: one lcs_box per process, held by the flow buffer
# ---------------------------------------------------------------------------
const contexts = Dict{Int,Ptr{Cvoid}}()

function context(device::Integer)
    get!(contexts, device) do
        ctx = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:lcs_ctx_create, libmi355x), Cint, (Cint, Ref{Ptr{Cvoid}}), device, ctx))
        ctx[]
    end
end

"""
    DeviceBox

Owns one `lcs_box` (fields, HSMAC partners, particle profiles in HBM); the
finalizer destroys it.
"""
mutable struct DeviceBox
    ptr::Ptr{Cvoid}
    nprofiles::Int
    function DeviceBox(ptr::Ptr{Cvoid})
        b = new(ptr, 0)
        live_boxes[b] = nothing
        finalizer(b) do x
            x.ptr == C_NULL || ccall((:lcs_box_destroy, libmi355x), Cint, (Ptr{Cvoid},), x.ptr)
            x.ptr = C_NULL
        end
    end
end
Base.unsafe_convert(::Type{Ptr{Cvoid}}, b::DeviceBox) = b.ptr

# the boxes KA.synchronize drains (weak: a collected box leaves by itself)
const live_boxes = WeakKeyDict{DeviceBox,Nothing}()

"""
    FieldRef(box, id)

A handle to one device field (`Us[1..3]`, `Us2[1..3]`, `P`); what the
reference's `Fields` tuples hold for this backend, so `synchalo!` calls on
`(buf.flow.fields.Us..., buf.flow.fields.P)` keep their shape.
"""
struct FieldRef
    box::DeviceBox
    id::Cint
end

struct DeviceFields
    Us::NTuple{3,FieldRef}
    Us2::NTuple{3,FieldRef}
    P::FieldRef
end

"""
    FlowBuffer <: LCS.AbstractBuffer

`buf.flow` for this backend (`Flows.FlowBuffer`, `src/flows/buffer.jl:115-121`,
without host arrays): the device box and handles to its fields.
"""
struct FlowBuffer <: LCS.AbstractBuffer
    box::DeviceBox
    fields::DeviceFields
end

"""
    ParticleBuffer <: LCS.AbstractBuffer

`buf.particles[i]` for this backend (`Particles.ParticleBuffer`,
`src/particles/buffer.jl:124-135`): profile `profile` (0-based) of the box.
"""
struct ParticleBuffer <: LCS.AbstractBuffer
    box::DeviceBox
    profile::Cint
    nvalid::Int
    ntotal::Int
end

dtype_of(::Type{Float64}) = LCS_F64
dtype_of(::Type{Float32}) = LCS_F32

function make_box(config::MI355XConfig, topo::Topologies.Topology)
    Topologies.is_multi_processing(topo) &&
        throw(ArgumentError("the mi355x backend advances one box per process on one GPU; use proc_dims [1, 1, 1]"))
    g = config.grid
    p = config.flow.params
    desc = Ref(BoxDesc(g.dims..., g.halo_size, dtype_of(LCS.FP), p.U0_m_s, p.L0_m, p.NU_m2_s))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    device = parse(Int, get(ENV, "LCS_MI355X_DEVICE", "0"))
    check(ccall((:lcs_box_create, libmi355x), Cint, (Ptr{Cvoid}, Ref{BoxDesc}, Ref{Ptr{Cvoid}}), context(device),
                desc, out))
    DeviceBox(out[])
end

drag_id(::Particles.LinearDrag) = Cint(0)
drag_id(::Particles.NonlinearDrag) = Cint(1)

function add_profile(box::DeviceBox, config::MI355XConfig, topo::Topologies.Topology, i::Integer)
    pc = config.particles[i]
    nvalid = Particles.nvalid_l(pc.population, topo)
    ntotal = Particles.ntotal_l(pc.population, topo)
    d = Ref(ParticleDesc(nvalid, ntotal, pc.params.gravity, pc.params.density_ratio, drag_id(pc.drag),
                         pc.cell.grid_ratio))
    idx = Ref{Cint}(0)
    check(ccall((:lcs_particles_add, libmi355x), Cint, (Ptr{Cvoid}, Ref{ParticleDesc}, Ref{Cint}), box, d, idx))
    box.nprofiles += 1
    ParticleBuffer(box, idx[], nvalid, ntotal)
end

function flow_buffer(box::DeviceBox)
    f(id) = FieldRef(box, id)
    FlowBuffer(box, DeviceFields((f(FIELD_U), f(FIELD_V), f(FIELD_W)), (f(FIELD_U2), f(FIELD_V2), f(FIELD_W2)),
                                 f(FIELD_P)))
end

# LCS.Buffer (src/buffer.jl:27-34): halo buffers are not needed on one
# process (SPHaloBuffer, lib/topologies/src/halo.jl:16,41)
function LCS.Buffer(::LCS.FlowMode, config::MI355XConfig, topo::Topologies.Topology)
    LCS.Buffer(Topologies.SPHaloBuffer(), flow_buffer(make_box(config, topo)), nothing)
end

function LCS.Buffer(::LCS.FlowParticleMode, config::MI355XConfig, topo::Topologies.Topology)
    box = make_box(config, topo)
    particles = if isnothing(config.particles)
        nothing
    else
        ntuple(i -> add_profile(box, config, topo, i), length(config.particles))
    end
    LCS.Buffer(Topologies.SPHaloBuffer(), flow_buffer(box), particles)
end

box_of(buf::LCS.Buffer) = buf.flow.box

# ---------------------------------------------------------------------------
# Initial conditions: the reference's own init! on host arrays, then upload
# ---------------------------------------------------------------------------
function upload(box::DeviceBox, field::Cint, A::AbstractArray)
    H = Array{LCS.FP}(parent(A))
    check(ccall((:lcs_field_upload, libmi355x), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Csize_t), box, field, H,
                sizeof(H)))
end

function upload_prop(pb::ParticleBuffer, prop::Integer, A::AbstractVector)
    H = Array(A)
    length(H) == pb.ntotal || throw(ArgumentError("property $prop: expected $(pb.ntotal) elements, got $(length(H))"))
    check(ccall((:lcs_particles_upload, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Csize_t), pb.box,
                pb.profile, prop, H, sizeof(H)))
end

# init! (src/init.jl:6-9): every flow and particle init kind of the reference,
# restart included, runs unchanged on a host copy of the buffers
function LCS.init!(buf::LCS.Buffer, state::LCS.State, config::MI355XConfig, topo::Topologies.Topology)
    hconfig = @set config.backend = KA.CPU()
    hbuf = LCS.Buffer(hconfig, topo)
    LCS.init!(hbuf, state, hconfig, topo)
    box = box_of(buf)
    f = hbuf.flow.fields
    for (id, A) in zip((FIELD_U, FIELD_V, FIELD_W, FIELD_U2, FIELD_V2, FIELD_W2, FIELD_P), (f.Us..., f.Us2..., f.P))
        upload(box, id, A)
    end
    if !isnothing(buf.particles)
        for (i, pb) in enumerate(buf.particles)
            nv = state.particles[i].nvalid
            nv == pb.nvalid || throw(ArgumentError(
                "profile $i: init left $nv valid particles, the device profile holds $(pb.nvalid) " *
                "(a profile's population is fixed when it is created)"))
            for (q, A) in enumerate(Tuple(hbuf.particles[i].props))  # Tuple(::Properties) order, ids first
                upload_prop(pb, q - 1, A)
            end
        end
    end
    KA.synchronize(config.backend)
    nothing
end

# ---------------------------------------------------------------------------
# Phases
# ---------------------------------------------------------------------------
stage_no(::LCS.RKStage1) = Cint(1)
stage_no(::LCS.RKStage2) = Cint(2)

# KA.synchronize (called by @log/@profile, src/log.jl:12,33,36): drain the
# stream of every live box of this process
function KA.synchronize(::MI355XBackend)
    for b in keys(live_boxes)
        b.ptr == C_NULL || check(ccall((:lcs_sync, libmi355x), Cint, (Ptr{Cvoid},), b))
    end
    nothing
end

# Topologies.synchalo! single process (src/topologies.jl:36-45,
# lib/topologies/src/halo.jl:138-159): A[i] = A[mod1(i, n)] on the ghosts
function Topologies.synchalo!(style::Topologies.HaloStyle, As::Tuple{Vararg{FieldRef}}, ::Topologies.HaloBuffer,
                              ::LCS.Grid, ::MI355XBackend, ::Topologies.Topology)
    isempty(As) && return nothing
    mask = UInt32(0)
    for A in As
        A.box === As[1].box || throw(ArgumentError("synchalo!: fields of different boxes"))
        mask |= UInt32(1) << A.id
    end
    full = style isa Topologies.FullHalo ? Cint(1) : Cint(0)
    check(ccall((:lcs_fill_ghost, libmi355x), Cint, (Ptr{Cvoid}, Cuint, Cint, Cint), As[1].box, mask, style.size,
                full))
end

# Flows.timestep (src/flows/timestep.jl:6-65)
function Flows.timestep(buf::LCS.Buffer, ::LCS.State, config::MI355XConfig, ::Topologies.Topology)
    dt = Ref{Cdouble}(0)
    check(ccall((:lcs_flow_timestep, libmi355x), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Cdouble}, Ptr{Cdouble}),
                box_of(buf), config.timestep.cfl, config.timestep.dtmax, dt, C_NULL))
    dt[]
end

# Particles.timestep (src/particles/timestep.jl:6-13)
function Particles.timestep(buf::LCS.Buffer, ::LCS.State, config::MI355XConfig, ::Topologies.Topology)
    isnothing(buf.particles) && return config.timestep.dtmax
    minimum(buf.particles) do pb
        dt = Ref{Cdouble}(0)
        check(ccall((:lcs_particles_timestep, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cdouble, Ref{Cdouble}), pb.box,
                    pb.profile, config.timestep.dtmax, dt))
        dt[]
    end
end

# Flows.rhs! (src/flows/rhs.jl:9-19): dU/dt of Us (stage 1) or Us2 (stage 2)
function Flows.rhs!(buf::LCS.Buffer, ::LCS.State, ::MI355XConfig, ::Topologies.Topology, stage::LCS.RKStage)
    check(ccall((:lcs_rhs, libmi355x), Cint, (Ptr{Cvoid}, Cint), box_of(buf), stage_no(stage)))
    LCS.EMPTY_LOG
end

# Flows.update! (src/flows/update.jl:12-22)
function Flows.update!(buf::LCS.Buffer, state::LCS.State, ::MI355XConfig, ::Topologies.Topology, stage::LCS.RKStage)
    check(ccall((:lcs_update, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cdouble), box_of(buf), stage_no(stage),
                LCS.dtrk(stage, state)))
    LCS.EMPTY_LOG
end

function coloring_id(c)
    c isa Parallel.RedBlackFast && return Cint(0)
    c isa Parallel.RedBlack && return Cint(1)
    throw(ArgumentError("the mi355x backend runs HSMAC with red-black-fast or red-black colouring, not $(typeof(c))"))
end

hsmac_params(c::Flows.HSMAC) = HsmacParams(c.time_blocking, c.overlap, coloring_id(c.coloring), c.omega, c.epsp,
                                           c.itrmin, c.itrmax)

# Flows.couple!(::HSMAC) (src/flows/couples/couple.jl:13-17, hsmac.jl:52-201)
function Flows.couple!(c::Flows.HSMAC, buf::LCS.Buffer, state::LCS.State, ::MI355XConfig, ::Topologies.Topology,
                       stage::LCS.RKStage)
    prm = Ref(hsmac_params(c))
    itr = Ref{Cint}(0)
    div = Ref{Cdouble}(0)
    check(ccall((:lcs_hsmac, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cdouble, Ref{HsmacParams}, Ref{Cint}, Ref{Cdouble}),
                box_of(buf), stage_no(stage), LCS.dtrk(stage, state), prm, itr, div))
    (; itr=Int(itr[]), div=div[])
end

# Flows.forcing!(::ConstantPowerLF) (src/flows/forcings/lf.jl:6-25)
function Flows.forcing!(f::Flows.ConstantPowerLF, buf::LCS.Buffer, state::LCS.State, ::MI355XConfig,
                        ::Topologies.Topology, stage::LCS.RKStage2)
    p = Ref{Cdouble}(0)
    check(ccall((:lcs_force_linear, libmi355x), Cint, (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Cdouble}), box_of(buf),
                f.power, LCS.dtrk(stage, state), p))
    (; p=p[])
end

# Flows.forcing!(::EnergyPreserveRCF) (src/flows/forcings/rcf.jl:1-100): box
# mean, spectrum rescaling and sparse inverse DFT all on the device
function Flows.forcing!(f::Flows.EnergyPreserveRCF, buf::LCS.Buffer, state::LCS.State, ::MI355XConfig,
                        ::Topologies.Topology, stage::LCS.RKStage2)
    p = Ref{Cdouble}(0)
    check(ccall((:lcs_force_rcf, libmi355x), Cint, (Ptr{Cvoid}, Cdouble, Cint, Cint, Cint, Cdouble, Ref{Cdouble}),
                box_of(buf), f.energy, f.kmin, f.kmax, f.coarse_size, LCS.dtrk(stage, state), p))
    (; p=p[], e10=f.energy)
end

# Particles.boundary! (src/particles/boundary.jl:33-59): one process, no exchange
function Particles.boundary!(buf::LCS.Buffer, state::LCS.State, ::MI355XConfig, topo::Topologies.Topology,
                             stage::LCS.RKStage, iprofile::Integer)
    pb = buf.particles[iprofile]
    check(ccall((:lcs_particles_boundary, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cint), pb.box, pb.profile,
                stage_no(stage)))
    nvalid = state.particles[iprofile].nvalid
    (; nvalid=Particles.ExchangeSummary(nvalid, topo), nrecvs=Particles.ExchangeSummary(0, topo),
     nsends=Particles.ExchangeSummary(0, topo))
end

# Particles.motion! (src/particles/motion.jl:15-36)
function Particles.motion!(buf::LCS.Buffer, ::LCS.State, ::MI355XConfig, ::Topologies.Topology, stage::LCS.RKStage,
                           iprofile::Integer)
    pb = buf.particles[iprofile]
    check(ccall((:lcs_particles_motion, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cint), pb.box, pb.profile,
                stage_no(stage)))
    LCS.EMPTY_LOG
end

# Particles.update! (src/particles/update.jl:18-31)
function Particles.update!(buf::LCS.Buffer, state::LCS.State, ::MI355XConfig, ::Topologies.Topology,
                           stage::LCS.RKStage, iprofile::Integer)
    pb = buf.particles[iprofile]
    check(ccall((:lcs_particles_update, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cint, Cdouble), pb.box, pb.profile,
                stage_no(stage), LCS.dtrk(stage, state)))
    LCS.EMPTY_LOG
end

# Particles.makeindex! (src/particles/makeindex.jl:9-27)
function Particles.makeindex!(buf::LCS.Buffer, ::LCS.State, ::MI355XConfig, ::Topologies.Topology,
                              stage::LCS.RKStage, iprofile::Integer)
    pb = buf.particles[iprofile]
    check(ccall((:lcs_particles_makeindex, libmi355x), Cint, (Ptr{Cvoid}, Cint, Cint), pb.box, pb.profile,
                stage_no(stage)))
    LCS.EMPTY_LOG
end

# ---------------------------------------------------------------------------
# The whole step: one C call, the reference's log shape
# ---------------------------------------------------------------------------
function forcing_desc(f)
    f isa Flows.ConstantPowerLF && return ForcingDesc(FORCING_LINEAR_POWER, f.power, 0.0, 0, 0, 0)
    f isa Flows.EnergyPreserveRCF && return ForcingDesc(FORCING_RC_ENERGY, 0.0, f.energy, f.kmin, f.kmax, f.coarse_size)
    f isa Flows.NoForcing && return ForcingDesc(FORCING_NONE, 0.0, 0.0, 0, 0, 0)
    throw(ArgumentError("the mi355x backend has no forcing $(typeof(f))"))
end

"""
    LOCALIZE_INTERVAL

Steps between reorderings of particle storage by position (`lcs_particles_localize`;
results do not change).  `LCS_MI355X_LOCALIZE` overrides it; 0 never reorders.
"""
const LOCALIZE_INTERVAL = Ref(parse(Int, get(ENV, "LCS_MI355X_LOCALIZE", "30")))

step_config(config::MI355XConfig) = StepConfig(config.timestep.cfl, config.timestep.dtmax,
                                               hsmac_params(config.flow.couple), forcing_desc(config.flow.forcing),
                                               LOCALIZE_INTERVAL[])

forcing_log(::Flows.NoForcing, l::StepLog) = LCS.EMPTY_LOG
forcing_log(::Flows.ConstantPowerLF, l::StepLog) = (; p=l.p)
forcing_log(::Flows.EnergyPreserveRCF, l::StepLog) = (; p=l.p, e10=l.e10)

# timestep! (src/timestep.jl:25-27).  lcs_step computes the step's dt itself,
# on the device, at its start (the same CFL and particle dt,
# src/timestep.jl:14-18), and evolve! stores it in state.dt; computing it
# here as well would run the CFL pass twice a step.  So timestep! leaves
# state.dt to evolve!, and the phase path (evolve_phases!) computes it first.
LCS.timestep!(::LCS.Buffer, state::LCS.State, ::MI355XConfig, ::Topologies.Topology) = state.dt

function run_step!(buf::LCS.Buffer, state::LCS.State, config::MI355XConfig)
    cfg = Ref(step_config(config))
    log = Ref(StepLog())
    check(ccall((:lcs_step, libmi355x), Cint, (Ptr{Cvoid}, Ref{StepConfig}, Ref{StepLog}), box_of(buf), cfg, log))
    l = log[]
    # evolve!(state) (src/evolve.jl:10-13); the device advanced its own copy
    state.dt = l.dt
    state.step += 1
    state.t += l.dt
    l
end

flow_log(config, l::StepLog) = (;
    rhs1=LCS.EMPTY_LOG, update1=LCS.EMPTY_LOG, couple1=(; itr=Int(l.itr1), div=l.div1),
    rhs2=LCS.EMPTY_LOG, update2=LCS.EMPTY_LOG, forcing=forcing_log(config.flow.forcing, l),
    couple2=(; itr=Int(l.itr2), div=l.div2))

# evolve!(::FlowMode) (src/evolve.jl:15-61)
function LCS.evolve!(::LCS.FlowMode, buf::LCS.Buffer, state::LCS.State, config::MI355XConfig,
                     ::Topologies.Topology)
    l = run_step!(buf, state, config)
    (; flow=flow_log(config, l))
end

# evolve!(::FlowParticleMode) (src/evolve.jl:63-146).  A callback between the
# RK stages needs the phases one by one, so it takes the phase path.
function LCS.evolve!(mode::LCS.FlowParticleMode, buf::LCS.Buffer, state::LCS.State, config::MI355XConfig,
                     topo::Topologies.Topology; cb_between_rk_stages=nothing)
    isnothing(cb_between_rk_stages) || return evolve_phases!(mode, buf, state, config, topo, cb_between_rk_stages)
    l = run_step!(buf, state, config)
    np = isnothing(buf.particles) ? 0 : length(buf.particles)
    ex = ntuple(i -> (; nvalid=Particles.ExchangeSummary(state.particles[i].nvalid, topo),
                      nrecvs=Particles.ExchangeSummary(0, topo), nsends=Particles.ExchangeSummary(0, topo)), np)
    empty = ntuple(_ -> LCS.EMPTY_LOG, np)
    (; particles=(; boundary1=ex, motion1=empty, update1=empty, boundary2=ex, motion2=empty, update2=empty),
     flow=flow_log(config, l))
end

# the reference's phase order (src/evolve.jl:63-146) through the phase overrides
function evolve_phases!(::LCS.FlowParticleMode, buf, state, config, topo, cb)
    state.dt = LCS.timestep(buf, state, config, topo)
    f = buf.flow.fields
    np = isnothing(buf.particles) ? 0 : length(buf.particles)
    stage = LCS.RKStage1()
    Topologies.synchalo!(Topologies.FullHalo(config.grid), (f.Us..., f.P), buf.halo, config.grid, config.backend, topo)
    pb1 = ntuple(i -> Particles.boundary!(buf, state, config, topo, stage, i), np)
    pm1 = ntuple(i -> Particles.motion!(buf, state, config, topo, stage, i), np)
    pu1 = ntuple(i -> Particles.update!(buf, state, config, topo, stage, i), np)
    r1 = Flows.rhs!(buf, state, config, topo, stage)
    u1 = Flows.update!(buf, state, config, topo, stage)
    c1 = Flows.couple!(config.flow.couple, buf, state, config, topo, stage)
    cb()
    stage = LCS.RKStage2()
    Topologies.synchalo!(Topologies.FullHalo(config.grid), (f.Us2..., f.P), buf.halo, config.grid, config.backend,
                         topo)
    pb2 = ntuple(i -> Particles.boundary!(buf, state, config, topo, stage, i), np)
    pm2 = ntuple(i -> Particles.motion!(buf, state, config, topo, stage, i), np)
    pu2 = ntuple(i -> Particles.update!(buf, state, config, topo, stage, i), np)
    r2 = Flows.rhs!(buf, state, config, topo, stage)
    u2 = Flows.update!(buf, state, config, topo, stage)
    fo = Flows.forcing!(config.flow.forcing, buf, state, config, topo, stage)
    c2 = Flows.couple!(config.flow.couple, buf, state, config, topo, stage)
    state.step += 1
    state.t += state.dt
    # keep the device's step counter in line with the host state
    check(ccall((:lcs_set_state, libmi355x), Cint, (Ptr{Cvoid}, Int64, Cdouble), box_of(buf), state.step, state.t))
    (; particles=(; boundary1=pb1, motion1=pm1, update1=pu1, boundary2=pb2, motion2=pm2, update2=pu2),
     flow=(; rhs1=r1, update1=u1, couple1=c1, rhs2=r2, update2=u2, forcing=fo, couple2=c2))
end

Flows.forcing!(::Flows.NoForcing, ::LCS.Buffer, ::LCS.State, ::MI355XConfig, ::Topologies.Topology,
               ::LCS.RKStage2) = LCS.EMPTY_LOG

# ---------------------------------------------------------------------------
# Statistics and restart files (cold paths)
# ---------------------------------------------------------------------------
summary(mean, max, min) = LCS.Summary{Float64}(mean, max, min)

# Flows.stat (src/flows/stat.jl:54-155)
function Flows.stat(buf::LCS.Buffer, ::LCS.State, config::MI355XConfig, ::Topologies.Topology)
    st = Ref{FlowStats}()
    check(ccall((:lcs_flow_stat, libmi355x), Cint, (Ptr{Cvoid}, Cint, Ref{FlowStats}), box_of(buf),
                config.flow.stat.integral_length, st))
    s = st[]
    Flows.FlowStat(; u=summary(s.u_mean, s.u_max, s.u_min), v=summary(s.v_mean, s.v_max, s.v_min),
                   w=summary(s.w_mean, s.w_max, s.w_min), k=s.k, ϵ=s.eps, τη=s.tau_eta, lμ=s.l_mu, lη=s.l_eta,
                   Reλ=s.Re_lambda, urms=s.urms, kη=s.k_eta, skew=s.skew, flat=s.flat, L=s.L, Cϵ=s.C_eps)
end

const SPECTRUM_WAVENUMBERS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768)

# Particles.stat (src/particles/stat.jl:87-155)
function Particles.stat(fstat::Flows.FlowStat, buf::LCS.Buffer, ::LCS.State, config::MI355XConfig,
                        ::Topologies.Topology, iprofile::Integer)
    pb = buf.particles[iprofile]
    sc = config.particles[iprofile].stat
    prm = Ref(ParticleStatParams(sc.rdf.contact_tolerance, sc.rdf.bin_size, sc.rdf.min_radius,
                                 sc.rdf.search_cell_size, sc.spectrum.max_wavenumber, sc.spectrum.nsamples))
    nb = sc.rdf.bin_size
    nsh = sc.spectrum.max_wavenumber > 0 ? count(<=(sc.spectrum.max_wavenumber), SPECTRUM_WAVENUMBERS) : 0
    edges = zeros(nb + 1)
    npairs = zeros(Int64, nb)
    gr = zeros(nb)
    wn = zeros(max(nsh, 1))
    enp = zeros(max(nsh, 1))
    st = Ref{ParticleStats}()
    check(ccall((:lcs_particle_stat, libmi355x), Cint,
                (Ptr{Cvoid}, Cint, Ref{ParticleStatParams}, Cdouble, Ref{ParticleStats}, Ptr{Cdouble}, Ptr{Int64},
                 Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                pb.box, pb.profile, prm, fstat.τη, st, edges, npairs, gr, wn, enp))
    s = st[]
    rdf = Particles.RDFStat(; gr_contact=s.gr_contact, npairs_contact=Int(s.npairs_contact), edges, gr,
                            npairs=Vector{Int}(npairs))
    spectrum = Particles.SpectrumStat(; wavenumbers=wn[1:nsh], Enp=enp[1:nsh])
    Particles.ParticleStat(; x=summary(s.x_mean, s.x_max, s.x_min), y=summary(s.y_mean, s.y_max, s.y_min),
                           z=summary(s.z_mean, s.z_max, s.z_min), u=summary(s.u_mean, s.u_max, s.u_min),
                           v=summary(s.v_mean, s.v_max, s.v_min), w=summary(s.w_mean, s.w_max, s.w_min),
                           diam=summary(s.diam_mean, s.diam_max, s.diam_min), τp=s.taup, St=s.St, rdf, spectrum)
end

# LCSIO.saverestart (src/LCSIO/save.jl:45-56): the reference's file layout,
# written by liblcs_io.so straight from the device
function LCSIO.saverestart(file::String, buf::LCS.Buffer, ::LCS.State, ::MI355XConfig, ::Topologies.Topology)
    rc = ccall((:lcs_save_restart, libio), Cint, (Ptr{Cvoid}, Cstring), box_of(buf), file)
    rc == 0 || throw(ErrorException("lcs_save_restart($file) failed ($rc): " *
                                    unsafe_string(ccall((:lcs_io_last_error, libio), Cstring, ()))))
    nothing
end

end # module
