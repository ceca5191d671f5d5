# Parity of the mi355x backend with the reference's own CPU backend, for a
# maintainer with Julia and an MI355X: the same config, once with
# backend cpu and once with backend mi355x, stepped side by side.  Bar: the
# HSMAC sweep counts of every step equal, U, V, W, P within relative L2 1e-8
# after 20 steps (FP64), particle positions by id within 1e-10.
using Accessors
using KernelAbstractions
using LCS
using LCSMI355X
using Test
using Topologies

const KA = KernelAbstractions

relerr(a, b) = sqrt(sum(abs2, a .- b) / max(sum(abs2, b), eps()))

function download(buf, id, shape)
    A = Array{LCS.FP}(undef, shape)
    LCSMI355X.check(ccall((:lcs_field_download, LCSMI355X.libmi355x), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Csize_t),
                          buf.flow.box, id, A, sizeof(A)))
    A
end

function download_prop(pb, prop)
    A = Vector{prop == 0 ? Int64 : LCS.FP}(undef, pb.ntotal)
    LCSMI355X.check(ccall((:lcs_particles_download, LCSMI355X.libmi355x), Cint,
                          (Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}, Csize_t), pb.box, pb.profile, prop, A, sizeof(A)))
    A
end

@testset "mi355x vs cpu: precompile.lcs-yaml" begin
    file = joinpath(dirname(pathof(LCS)), "..", "precompile.lcs-yaml")
    cpu = LCSIO.readconfig(file)
    @reset cpu.backend = KA.CPU()
    @reset cpu.flow.stat.integral_length = false
    gpu = @set cpu.backend = LCSMI355X.MI355XBackend()
    topo = Topologies.Topology(cpu)
    bc, sc = LCS.Buffer(cpu, topo), LCS.State(cpu, topo)
    bg, sg = LCS.Buffer(gpu, topo), LCS.State(gpu, topo)
    LCS.init!(bc, sc, cpu, topo)
    LCS.init!(bg, sg, gpu, topo)
    for s in 1:20
        LCS.timestep!(bc, sc, cpu, topo)
        lc = LCS.evolve!(bc, sc, cpu, topo)
        LCS.timestep!(bg, sg, gpu, topo)
        lg = LCS.evolve!(bg, sg, gpu, topo)
        @test lg.flow.couple1.itr == lc.flow.couple1.itr
        @test lg.flow.couple2.itr == lc.flow.couple2.itr
        @test sg.dt ≈ sc.dt rtol = 1e-12
        @test keys(lg) == keys(lc) && keys(lg.flow) == keys(lc.flow)
    end
    @test sg.step == sc.step
    f = bc.flow.fields
    shape = size(parent(f.P))
    for (id, A) in zip(Cint.((0, 1, 2, 6)), (f.Us..., f.P))
        @test relerr(download(bg, id, shape), Array(parent(A))) < 1e-8
    end
    for (i, pb) in enumerate(bg.particles)
        nv = sc.particles[i].nvalid
        idc = Array(bc.particles[i].props.ids)[1:nv]
        idg = download_prop(pb, 0)[1:nv]
        oc, og = sortperm(idc), sortperm(idg)
        @test idc[oc] == idg[og]
        for (q, A) in zip(2:4, bc.particles[i].props.xss)
            @test relerr(download_prop(pb, q)[1:nv][og], Array(A)[1:nv][oc]) < 1e-10
        end
    end
    fs_c = LCS.Flows.stat(bc, sc, cpu, topo)
    fs_g = LCS.Flows.stat(bg, sg, gpu, topo)
    @test fs_g.k ≈ fs_c.k rtol = 1e-8
    @test fs_g.ϵ ≈ fs_c.ϵ rtol = 1e-8
end
