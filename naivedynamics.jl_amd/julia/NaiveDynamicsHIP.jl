# NaiveDynamicsHIP.jl -- Julia binding of libnaivehip.so for NaiveDynamics.jl.
#
# Adds methods to NaiveDynamics' entry points that dispatch on a `HIPBackend`
# argument and `ccall` the C ABI declared in include/naivedynamics.h.  The
# stub functions of src/PkgExtensions.jl:37-67 (gpubvh_neighborlist,
# simdbuild_traverse_bvh, ...) get their HIP methods here; the CPU entry points
# keep theirs.  Indices returned to Julia are 1-based as in the reference.
#
# Untested in this repository (no Julia toolchain in its build image); the
# Python host mirror in naivedynamics.jl_amd/ calls the same symbols and is
# the tested consumer.
module NaiveDynamicsHIP

using NaiveDynamics
using StaticArrays

export HIPBackend, cell_neighborlist

const libnaivehip = get(ENV, "NAIVEHIP_LIB", "libnaivehip")

struct HIPBackend
    handle::Ptr{Cvoid}
end

struct NDSpec                      # nd_bvh_spec == SpheresBVHSpecs{Float32,Int32}
    neighbor_distance::Float32
    atom_count::Int32
    leaves_count::Int32
    branches_count::Int32
    atomsperleaf::Int32
end

struct NDNode                      # nd_node == GridKey{Float32,Int32}
    min::SVector{3,Float32}
    max::SVector{3,Float32}
    left::Int32
    skip::Int32
end

struct NDSimParams                 # nd_sim_params
    n::Int32; precision::Int32; morton_mode::Int32; atomsperleaf::Int32
    rebuild_every::Int32; walls::Int32; energy_every::Int32; thermostat_every::Int32
    use_graph::Int32; max_neighbors::Int32; resort_every::Int32; list_walk::Int32
    pair_batch::Int32; walk_front::Int32; atom_lanes::Int32
    eps::Float64; sigma::Float64; rc::Float64; k_coul::Float64; rc_coul::Float64
    skin::Float64; dt::Float64; lo::Float64; hi::Float64; temperature::Float64; vdamp::Float64
end

const ND_ERR_CAPACITY = Cint(-3)

function HIPBackend(device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:nd_ctx_create, libnaivehip), Cint, (Int32, Ref{Ptr{Cvoid}}), device, h)
    rc == 0 || error("nd_ctx_create failed ($rc)")
    b = HIPBackend(h[])
    finalizer(x -> ccall((:nd_ctx_destroy, libnaivehip), Cint, (Ptr{Cvoid},), x.handle), Ref(b))
    return b
end

function check(b::HIPBackend, rc::Cint, what::AbstractString)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:nd_last_error, libnaivehip), Cstring, (Ptr{Cvoid},), b.handle))
    rc == Cint(-2) && throw(InexactError(Symbol(what), Int32, msg))   # reference: round(Int32, ...)
    error("$what: $msg (status $rc)")
end

ndspec(spec::SpheresBVHSpecs{Float32,Int32}) = NDSpec(spec.neighbor_distance, spec.atom_count,
    spec.leaves_count, spec.branches_count, spec.atomsperleaf)

# Vec3D (Vector{MVector{3,Float32}}) holds boxed objects; the C ABI needs 3 x N contiguous.
packxyz(position) = reduce(hcat, (Vector{Float32}(p) for p in position))::Matrix{Float32}

function pairlist(b::HIPBackend, position, spec::SpheresBVHSpecs{Float32,Int32}, predicate::Int32)
    xyz = packxyz(position)
    s = Ref(ndspec(spec))
    cap = Int64(48) * spec.atom_count
    while true
        out = Vector{NTuple{3,UInt32}}(undef, cap)          # nd_pair: 12 bytes
        count = Ref{Int64}(0)
        rc = ccall((:nd_leafbuild_traverse_bvh_host, libnaivehip), Cint,
                   (Ptr{Cvoid}, Ptr{Float32}, Ref{NDSpec}, Int32, Int32, Ptr{Cvoid}, Int64, Ref{Int64}),
                   b.handle, xyz, s, predicate, Int32(0), out, cap, count)
        if rc == ND_ERR_CAPACITY
            cap = count[] + count[] ÷ 8 + 1024
            continue
        end
        check(b, rc, "leafbuild_traverse_bvh")
        res = Vector{Tuple{Int32,Int32,Float32}}(undef, count[])
        for k in 1:count[]
            t = out[k]
            res[k] = (reinterpret(Int32, t[1]) + Int32(1), reinterpret(Int32, t[2]) + Int32(1),
                      reinterpret(Float32, t[3]))
        end
        return res
    end
end

# leafbuild_traverse_bvh(position, spec) (BVHTraverse.jl:1423-1428) on the GPU
NaiveDynamics.leafbuild_traverse_bvh(b::HIPBackend, position, spec) = pairlist(b, position, spec, Int32(0))

# simdbuild_traverse_bvh (ext/NaiveSIMD.jl:706-710): sqrt(s) < r predicate, apl == 4
function NaiveDynamics.simdbuild_traverse_bvh(b::HIPBackend, position, spec)
    @assert spec.atomsperleaf == 4 "This simd method requires 4 primitives per leaf"
    return pairlist(b, position, spec, Int32(1))
end

# device staging for the entry points that take device buffers
function devbuf(b::HIPBackend, bytes::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(b, ccall((:nd_malloc, libnaivehip), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}),
                   b.handle, max(bytes, 1), p), "nd_malloc")
    return p[]
end
devfree(b::HIPBackend, p) = ccall((:nd_free, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b.handle, p)
h2d(b::HIPBackend, d, h::Array) = check(b, ccall((:nd_memcpy_h2d, libnaivehip), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), b.handle, d, h, sizeof(h)), "nd_memcpy_h2d")
d2h(b::HIPBackend, h::Array, d) = check(b, ccall((:nd_memcpy_d2h, libnaivehip), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t), b.handle, h, d, sizeof(h)), "nd_memcpy_d2h")

const ND_QUERY_OWN_LEAF = Int32(0)
const ND_QUERY_ROOT = Int32(1)

# point-query list (nd_build_traverse_bvh) on a staged copy of position
function pointquery(b::HIPBackend, dxyz, spec::SpheresBVHSpecs{Float32,Int32}, query::Int32)
    s = Ref(ndspec(spec))
    cap = Int64(48) * spec.atom_count
    while true
        dout = devbuf(b, 12cap)
        count = Ref{Int64}(0)
        rc = ccall((:nd_build_traverse_bvh, libnaivehip), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Ref{NDSpec}, Int32, Int32, Ptr{Cvoid}, Int64, Ref{Int64}),
                   b.handle, dxyz, s, query, Int32(0), dout, cap, count)
        if rc == ND_ERR_CAPACITY
            devfree(b, dout)
            cap = count[] + count[] ÷ 8 + 1024
            continue
        end
        check(b, rc, "build_traverse_bvh")
        out = Vector{NTuple{3,UInt32}}(undef, count[])
        count[] > 0 && d2h(b, out, dout)
        devfree(b, dout)
        return [(reinterpret(Int32, t[1]) + Int32(1), reinterpret(Int32, t[2]) + Int32(1),
                 reinterpret(Float32, t[3])) for t in out]
    end
end

# tree, sorted primitives and store of the last one-call build (1-based)
function exporttree(b::HIPBackend, position, spec)
    nL = spec.leaves_count
    tree = Vector{NDNode}(undef, 2nL - 1)
    perm = Vector{Int32}(undef, spec.atom_count)
    codes = Vector{Int32}(undef, spec.atom_count)
    store = Vector{Int32}(undef, nL - 1)
    check(b, ccall((:nd_tree_export, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{NDNode}, Ptr{Int32}, Ptr{Int32}),
                   b.handle, tree, perm, codes), "tree export")
    check(b, ccall((:nd_tree_store_export, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Int32}), b.handle, store),
          "store export")
    keys = [GridKey{Float32,Int32}(n.min, n.max, n.left, n.skip) for n in tree]
    prim = APointPrimitive{Float32,Int32}(perm .+ Int32(1), codes, [SVector{3,Float32}(position[p + 1]) for p in perm])
    return keys, prim, perm .+ Int32(1), store
end

# build_traverse_bvh(position, spec) (BVHTraverse.jl:1416-1421): TreeData + neighbor_traverse
function NaiveDynamics.build_traverse_bvh(b::HIPBackend, position, spec)
    xyz = packxyz(position)
    dxyz = devbuf(b, sizeof(xyz))
    try
        h2d(b, dxyz, xyz)
        return pointquery(b, dxyz, spec, ND_QUERY_OWN_LEAF)
    finally
        devfree(b, dxyz)
    end
end

# gpubvh_neighborlist(backend, position, spec) (ext/NaiveKA.jl:470-558): root-started
# point queries over the +-r tree, with the whole treedata
function NaiveDynamics.gpubvh_neighborlist(b::HIPBackend, position, spec)
    xyz = packxyz(position)
    n = size(xyz, 2)
    dxyz = devbuf(b, sizeof(xyz))
    dq = devbuf(b, 12n)
    try
        h2d(b, dxyz, xyz)
        check(b, ccall((:nd_quantized_positions, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int32, Ptr{Cvoid}),
                       b.handle, dxyz, Int32(n), dq), "quantized_positions!")
        q = Matrix{Int32}(undef, 3, n)
        d2h(b, q, dq)
        pairs = pointquery(b, dxyz, spec, ND_QUERY_ROOT)
        keys, prim, _, store = exporttree(b, position, spec)
        quantized = [SVector{3,Int32}(q[1, i], q[2, i], q[3, i]) for i in 1:n]
        return (pairlist = pairs, treedata = (tree = keys, mortonpositions = prim,
                                              quantizedposition = quantized, store = store))
    finally
        devfree(b, dxyz)
        devfree(b, dq)
    end
end

# neighborlist(x, cutoff) of CellListMap (assets/BVHBenchSuite.jl:307, src/MDInput.jl:290):
# the uniform cell-list search, (i, j, d) with i < j, 1-based; sqrt(s) < cutoff
function cell_neighborlist(b::HIPBackend, position, cutoff::Real; predicate::Int32 = Int32(1))
    xyz = packxyz(position)
    n = Int32(size(xyz, 2))
    cap = Int64(48) * n + 1024
    while true
        out = Vector{NTuple{3,UInt32}}(undef, cap)
        count = Ref{Int64}(0)
        rc = ccall((:nd_cell_neighborlist_host, libnaivehip), Cint,
                   (Ptr{Cvoid}, Ptr{Float32}, Int32, Float32, Int32, Ptr{Cvoid}, Int64, Ref{Int64}),
                   b.handle, xyz, n, Float32(cutoff), predicate, out, cap, count)
        if rc == ND_ERR_CAPACITY
            cap = count[] + count[] ÷ 8 + 1024
            continue
        end
        check(b, rc, "neighborlist")
        return [(reinterpret(Int32, t[1]) + Int32(1), reinterpret(Int32, t[2]) + Int32(1),
                 reinterpret(Float32, t[3])) for t in view(out, 1:count[])]
    end
end

# force_lennardjones!(force, pairslist, position) (Forces.jl:15-45), textbook LJ
function NaiveDynamics.force_lennardjones!(b::HIPBackend, force, pairslist, position; eps = 1f0, sigma = 1f0, rc = 2.5f0)
    n = length(position)
    dx = devcopy(b, packxyz(position)); df = devalloc(b, 12n)
    dp = devcopy(b, [(p[1] - Int32(1), p[2] - Int32(1), Float32(p[3])) for p in pairslist])
    params = Ref((Float32(eps), Float32(sigma), Float32(rc), 0f0, 0f0, Int32(0), Int32(0)))
    e = Ref{Float64}(0)
    rc_ = ccall((:nd_force_lennardjones, libnaivehip), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ref{Float64}),
                b.handle, df, dp, length(pairslist), dx, n, params, e)
    check(b, rc_, "force_lennardjones!")
    f = Matrix{Float32}(undef, 3, n); download!(b, f, df)
    for i in 1:n; force[i] .= @view f[:, i]; end
    foreach(p -> devfree(b, p), (dx, df, dp))
    return e[]
end

# simulate!(sys, spec, clct) (Simulator.jl:154-256): duration steps on the GPU
function NaiveDynamics.simulate!(b::HIPBackend, sys, spec, clct; dt = 0.001, rc = 2.5, rc_coul = 2.5,
                                 skin = 0.3, rebuild_every = 10, precision = 0, lo = 0.0, hi = 1.0)
    n = length(sys.position)
    # key 2: Hilbert index (ND_MORTON_HILBERT), 16 atoms per leaf
    p = Ref(NDSimParams(n, precision, 2, 16, rebuild_every, 1, 10, 0, 1, 0, 0, 0, 0, 0, 0,
                        1.0, 1.0, rc, 1.0, rc_coul, skin, dt, lo, hi, 1.0, 1.0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(b, ccall((:nd_sim_create, libnaivehip), Cint, (Ptr{Cvoid}, Ref{NDSimParams}, Ref{Ptr{Cvoid}}), b.handle, p, h), "simulate!")
    x = Float64.(reduce(hcat, sys.position)); v = Float64.(reduce(hcat, sys.velocity))
    q = Float64.(sys.charge); m = Float64.(sys.mass)
    check(b, ccall((:nd_sim_set_state, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   h[], x, v, q, m), "simulate!")
    check(b, ccall((:nd_sim_run, libnaivehip), Cint, (Ptr{Cvoid}, Int32), h[], spec.duration), "simulate!")
    f = similar(x)
    check(b, ccall((:nd_sim_get_state, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   h[], x, v, f), "simulate!")
    ccall((:nd_sim_destroy, libnaivehip), Cint, (Ptr{Cvoid},), h[])
    for i in 1:n
        sys.position[i] .= @view x[:, i]; sys.velocity[i] .= @view v[:, i]; sys.force[i] .= @view f[:, i]
    end
    return [sys.position]
end

# --- device staging helpers (nd_malloc / nd_memcpy_*) ---
function devalloc(b::HIPBackend, bytes::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(b, ccall((:nd_malloc, libnaivehip), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}), b.handle, bytes, p), "nd_malloc")
    return p[]
end
function devcopy(b::HIPBackend, a::AbstractArray)
    p = devalloc(b, sizeof(a))
    check(b, ccall((:nd_memcpy_h2d, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                   b.handle, p, a, sizeof(a)), "nd_memcpy_h2d")
    return p
end
download!(b::HIPBackend, a::AbstractArray, p::Ptr{Cvoid}) =
    check(b, ccall((:nd_memcpy_d2h, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Csize_t),
                   b.handle, a, p, sizeof(a)), "nd_memcpy_d2h")
devfree(b::HIPBackend, p::Ptr{Cvoid}) = ccall((:nd_free, libnaivehip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), b.handle, p)

end # module
