"""
    MuscleHIPExt

Muscle.jl package extension that runs the contraction hot path on an AMD
MI355X through `libmuscle_hip.so` (a plain C ABI, `include/muscle_hip.h`).

It plugs in where `ext/MuscleCUDAExt.jl` plugs in for CUDA: a Domain for
`ROCArray`, rules that select `BackendHIP` for device operands, and the six
backend methods the front ends call:

  binary_einsum(::BackendHIP, inds_c, a, b)   (call site src/Operations/binary_einsum.jl:50)
  binary_einsum!(::BackendHIP, c, a, b)       (:68)
  hadamard(::BackendHIP, a, b)                (src/Operations/hadamard.jl:12)
  hadamard!(::BackendHIP, c, a, b)            (:27)
  unary_einsum(::BackendHIP, inds_y, x)       (src/Operations/unary_einsum.jl:35)
  unary_einsum!(::BackendHIP, y, x)           (:44)

and, for the two-site gate update that follows the contractions, the
factorization behind it:

  tensor_svd_thin(::BackendHIP, A; inds_u, inds_v, ind_s)   (src/Operations/tensor_svd.jl:59-68)
  tensor_svd_trunc(::BackendHIP, A; ..., threshold, maxdim)  (:160-215)
  tensor_qr_thin(::BackendHIP, A; inds_q, inds_r, ind_virtual) (src/Operations/tensor_qr.jl:27-68)
  simple_update(::BackendHIP, A, ...)                        (src/Operations/simple_update.jl:37-87)

It assumes `struct BackendHIP <: Backend end` next to the other backends
(src/Backend.jl:6-14) and `struct DomainROCm <: Domain end` next to the other
domains (src/Domain.jl:6-9).  Set ENV["MUSCLE_HIP_LIB"] to the library path.
"""
module MuscleHIPExt

using Muscle
using Muscle: Tensor, Index, inds, BackendHIP, DomainROCm
using AMDGPU

const LIB = get(ENV, "MUSCLE_HIP_LIB", "libmuscle_hip")

# --- C ABI mirror ------------------------------------------------------------

const DTYPE = Dict{DataType,Int32}(Float32 => 0, Float64 => 1, ComplexF32 => 2, ComplexF64 => 3)
const STATUS_OK, STATUS_INVALID, STATUS_DIM, STATUS_UNSUPPORTED = 0, 1, 2, 3

# muscle_hip_tensor_t
struct CTensor
    dtype::Int32
    rank::Int32
    data::Ptr{Cvoid}
    modes::Ptr{Int32}
    extents::Ptr{Int64}
    strides::Ptr{Int64}
end

# muscle_hip_info_t (read back for introspection)
struct CInfo
    kernel::Int32
    precision::Int32
    slices::Int32
    swapped::Int32
    m::Int64
    n::Int64
    k::Int64
    batch::Int64
    flops::Float64
    bytes::Float64
    err_bound::Float64
end

function check(status::Integer)
    status == STATUS_OK && return nothing
    msg = unsafe_string(ccall((:muscle_hip_last_error, LIB), Cstring, ()))
    status == STATUS_DIM && throw(DimensionMismatch(msg))
    status in (STATUS_INVALID, STATUS_UNSUPPORTED) && throw(ArgumentError(msg))
    name = unsafe_string(ccall((:muscle_hip_status_string, LIB), Cstring, (Cint,), status))
    error("libmuscle_hip: $name: $msg")
end

# One library handle per task and device; the stream is AMDGPU's task-local
# stream, set before every call (binary_einsum's backend override is
# task-scoped too, src/Backend.jl:16-18).
function handle()
    key = (:muscle_hip_handle, AMDGPU.device_id(AMDGPU.device()))
    h = get(task_local_storage(), key, C_NULL)
    if h == C_NULL
        ref = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:muscle_hip_create, LIB), Cint, (Ref{Ptr{Cvoid}},), ref))
        h = ref[]
        task_local_storage(key, h)
    end
    check(ccall((:muscle_hip_set_stream, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), h, AMDGPU.stream().stream))
    return h
end

# Mode labels: Dict(ind => i for (i, ind) in enumerate(unique(...))), as
# ext/MuscleCUDAExt.jl:24-27 builds them.
modemap(tensors...) = Dict{Index,Int32}(ind => Int32(i) for (i, ind) in enumerate(unique(vcat(map(inds, tensors)...))))

# A C descriptor of `t` with modes `labels`; the returned arrays must be kept
# alive (GC.@preserve) across the ccall.
function describe(t::Tensor, labels::Dict{Index,Int32})
    arr = parent(t)
    modes = Int32[labels[i] for i in inds(t)]
    exts = Int64[size(arr)...]
    strs = Int64[strides(arr)...]
    desc = CTensor(DTYPE[eltype(arr)], Int32(ndims(arr)), Ptr{Cvoid}(pointer(arr)),
                   pointer(modes), pointer(exts), pointer(strs))
    return desc, (modes, exts, strs)
end

function contract!(c::Tensor, a::Tensor, b::Tensor; precision::Integer=0, target::Real=1e-12)
    labels = modemap(a, b, c)
    dc, kc = describe(c, labels)
    da, ka = describe(a, labels)
    db, kb = describe(b, labels)
    GC.@preserve c a b kc ka kb begin
        check(ccall((:muscle_hip_binary_einsum, LIB), Cint,
                    (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}, Cint, Cdouble, Ptr{CInfo}),
                    handle(), dc, da, db, precision, target, C_NULL))
    end
    return c
end

# --- dispatch ------------------------------------------------------------------

Muscle.Domain(::Type{<:ROCArray}) = DomainROCm()

Muscle.choose_backend_rule(::typeof(binary_einsum), ::DomainROCm, ::DomainROCm) = BackendHIP()
Muscle.choose_backend_rule(::typeof(binary_einsum!), ::DomainROCm, ::DomainROCm, ::DomainROCm) = BackendHIP()
# more specific than the generic `::Domain` rules of src/Operations/hadamard.jl:4-5
Muscle.choose_backend_rule(::typeof(hadamard), ::DomainROCm, ::DomainROCm) = BackendHIP()
Muscle.choose_backend_rule(::typeof(hadamard!), ::DomainROCm, ::DomainROCm, ::DomainROCm) = BackendHIP()
Muscle.choose_backend_rule(::typeof(unary_einsum), ::DomainROCm) = BackendHIP()
Muscle.choose_backend_rule(::typeof(unary_einsum!), ::DomainROCm, ::DomainROCm) = BackendHIP()

function Muscle.binary_einsum(::BackendHIP, inds_c, a::Tensor, b::Tensor)
    sizes = Dict{Index,Int}()
    for t in (a, b), (i, n) in zip(inds(t), size(parent(t)))
        sizes[i] = n
    end
    T = Base.promote_eltype(parent(a), parent(b))
    c = Tensor(similar(parent(a), T, Tuple(sizes[i] for i in inds_c)), collect(inds_c))
    return contract!(c, a, b)
end

Muscle.binary_einsum!(::BackendHIP, c::Tensor, a::Tensor, b::Tensor) = contract!(c, a, b)

function Muscle.hadamard!(::BackendHIP, c::Tensor, a::Tensor, b::Tensor)
    labels = modemap(a, b)
    dc, kc = describe(c, labels)
    da, ka = describe(a, labels)
    db, kb = describe(b, labels)
    GC.@preserve c a b kc ka kb begin
        check(ccall((:muscle_hip_hadamard, LIB), Cint,
                    (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}, Ptr{CInfo}),
                    handle(), dc, da, db, C_NULL))
    end
    return c
end

# the result has a's indices and a's element type (copy(a), src/Operations/hadamard.jl:38)
Muscle.hadamard(::BackendHIP, a::Tensor, b::Tensor) =
    Muscle.hadamard!(BackendHIP(), Tensor(similar(parent(a)), inds(a)), a, b)

function Muscle.unary_einsum!(::BackendHIP, y::Tensor, x::Tensor)
    labels = modemap(x)
    dy, ky = describe(y, labels)
    dx, kx = describe(x, labels)
    GC.@preserve y x ky kx begin
        check(ccall((:muscle_hip_unary_einsum, LIB), Cint, (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ptr{CInfo}),
                    handle(), dy, dx, C_NULL))
    end
    return y
end

function Muscle.unary_einsum(::BackendHIP, inds_y, x::Tensor)
    sizes = Dict(zip(inds(x), size(parent(x))))
    y = Tensor(similar(parent(x), Tuple(sizes[i] for i in inds_y)), collect(inds_y))
    return Muscle.unary_einsum!(BackendHIP(), y, x)
end

# --- factorization ---------------------------------------------------------------

Muscle.choose_backend_rule(::typeof(Muscle.tensor_svd_thin), ::DomainROCm) = BackendHIP()
Muscle.choose_backend_rule(::typeof(Muscle.simple_update), ::DomainROCm, ::DomainROCm, ::DomainROCm) = BackendHIP()

# U[inds_u..., ind_s], s[ind_s], Vt[inds_v..., ind_s] with A = U s Vt, the
# layout of BackendBase's tensor_svd_thin (src/Operations/tensor_svd.jl:100-125)
function Muscle.tensor_svd_thin(::BackendHIP, A::Tensor; inds_u=(), inds_v=(), ind_s=Index(gensym(:vind)), kwargs...)
    inds_u, inds_v = Muscle.factorinds(inds(A), inds_u, inds_v)
    isdisjoint(inds_u, inds_v) || throw(ArgumentError("inds_u and inds_v must be disjoint"))
    issetequal(inds_u ∪ inds_v, inds(A)) || throw(ArgumentError("inds_u and inds_v must cover inds(A)"))
    ind_s ∉ inds(A) || throw(ArgumentError("ind_s is already an index of A"))
    su = [size(A, i) for i in inds_u]
    sv = [size(A, i) for i in inds_v]
    k = min(prod(su), prod(sv))
    T = eltype(parent(A))
    U = Tensor(similar(parent(A), T, su..., k), [inds_u; ind_s])
    s = Tensor(similar(parent(A), real(T), k), [ind_s])
    Vt = Tensor(similar(parent(A), T, sv..., k), [inds_v; ind_s])
    labels = modemap(U, s, Vt, A)
    du, ku = describe(U, labels)
    ds, ks = describe(s, labels)
    dv, kv = describe(Vt, labels)
    da, ka = describe(A, labels)
    GC.@preserve U s Vt A ku ks kv ka begin
        check(ccall((:muscle_hip_svd_thin, LIB), Cint,
                    (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}),
                    handle(), du, ds, dv, da))
    end
    return U, s, Vt
end

Muscle.choose_backend_rule(::typeof(Muscle.tensor_qr_thin), ::DomainROCm) = BackendHIP()

# Q[inds_q..., ind_virtual], R[ind_virtual, inds_r...] with A = Q R
# (src/Operations/tensor_qr.jl:45-68)
function Muscle.tensor_qr_thin(::BackendHIP, A::Tensor; inds_q=(), inds_r=(), ind_virtual=Index(gensym(:qr)),
                               kwargs...)
    ind_virtual ∉ inds(A) || throw(ArgumentError("new virtual bond name ($ind_virtual) cannot be already be present"))
    inds_q, inds_r = Muscle.factorinds(inds(A), inds_q, inds_r)
    issetequal(inds_q ∪ inds_r, inds(A)) || throw(ArgumentError("inds_q and inds_r must cover inds(A)"))
    sq = [size(A, i) for i in inds_q]
    sr = [size(A, i) for i in inds_r]
    k = min(prod(sq), prod(sr))
    Q = Tensor(similar(parent(A), sq..., k), [inds_q; ind_virtual])
    R = Tensor(similar(parent(A), k, sr...), [ind_virtual; inds_r])
    labels = modemap(Q, R, A)
    dq, kq = describe(Q, labels)
    dr, kr = describe(R, labels)
    da, ka = describe(A, labels)
    GC.@preserve Q R A kq kr ka begin
        check(ccall((:muscle_hip_qr_thin, LIB), Cint, (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}),
                    handle(), dq, dr, da))
    end
    return Q, R
end

# the thin SVD, then the cut of src/Operations/tensor_svd.jl:183-203 on the host
function Muscle.tensor_svd_trunc(::BackendHIP, A::Tensor; inds_u=(), inds_v=(), ind_s=Index(gensym(:vind)),
                                 threshold=nothing, maxdim=nothing, kwargs...)
    U, s, Vt = Muscle.tensor_svd_thin(BackendHIP(), A; inds_u, inds_v, ind_s)
    sh = Array(parent(s))
    k = length(sh)
    isnothing(maxdim) || (k = min(k, maxdim))
    isnothing(threshold) || (k = something(findfirst(<(sqrt(sum(abs2, sh)) * threshold), view(sh, 1:k)), k))
    return view(U, ind_s => 1:k), view(s, ind_s => 1:k), view(Vt, ind_s => 1:k)
end

# simple_update.jl:37-87 on device tensors: the contractions and the absorption
# run through this backend's binary_einsum and hadamard!, the SVD through
# muscle_hip_svd_thin
function Muscle.simple_update(::BackendHIP, A::Tensor, ind_physical_a::Index, B::Tensor, ind_physical_b::Index,
                              ind_bond_ab::Index, G::Tensor, ind_physical_g_a::Index, ind_physical_g_b::Index;
                              normalize::Bool=false, absorb::Muscle.AbsorbBehavior=Muscle.DontAbsorb(),
                              atol::Float64=0.0, rtol::Float64=0.0, maxdim=nothing)
    Θ = Muscle.binary_einsum(Muscle.binary_einsum(A, B; dims=[ind_bond_ab]), G; dims=[ind_physical_a, ind_physical_b])
    Θ = replace(Θ, ind_physical_g_a => ind_physical_a, ind_physical_g_b => ind_physical_b)
    U, S, V = Muscle.tensor_svd_thin(BackendHIP(), Θ; inds_u=setdiff(inds(A), [ind_bond_ab]),
                                     inds_v=setdiff(inds(B), [ind_bond_ab]), ind_s=ind_bond_ab)
    if !isnothing(maxdim)
        χ = min(maxdim, length(S))
        U, S, V = view(U, ind_bond_ab => 1:χ), view(S, ind_bond_ab => 1:χ), view(V, ind_bond_ab => 1:χ)
    end
    normalize && (parent(S) ./= sqrt(sum(abs2, parent(S))))
    absorb isa Muscle.DontAbsorb && return U, S, V
    if absorb isa Muscle.AbsorbU
        U = Muscle.hadamard!(U, U, S)
    elseif absorb isa Muscle.AbsorbV
        V = Muscle.hadamard!(V, V, S)
    elseif absorb isa Muscle.AbsorbEqually
        S_sqrt = Tensor(sqrt.(parent(S)), inds(S))
        U = Muscle.hadamard!(U, U, S_sqrt)
        V = Muscle.hadamard!(V, V, S_sqrt)
    end
    return U, V
end

end # module
