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

and the factorizations of the two-site gate update (whose generic method,
src/Operations/simple_update.jl:35-82, runs on device tensors once the rule
below selects this backend):

  tensor_svd_thin(::BackendHIP, A; inds_u, inds_v, ind_s)      (src/Operations/tensor_svd.jl:59-68,100-125)
  tensor_svd_thin!(::BackendHIP, U, s, V, A)                    (:72-79,129)
  tensor_svd_trunc(::BackendHIP, A; ..., threshold, maxdim)     (:81-95,160-215)
  tensor_svd_trunc!(::BackendHIP, U, s, V, A; ...)              (:217)
  tensor_qr_thin(::BackendHIP, A; inds_q, inds_r, ind_virtual)  (src/Operations/tensor_qr.jl:27-68)
  tensor_qr_thin!(::BackendHIP, Q, R, A)                        (:36-43,70)

`BackendHIP` carries the arithmetic of binary_einsum, which the front end
cannot pass as keyword arguments (the CUDA extension forwards its kwargs to
cuTENSOR instead, ext/MuscleCUDAExt.jl:22,37):

  with_backend(BackendHIP(precision=:ozaki_int8, target_rel_error=1e-10)) do
      binary_einsum(a, b)
  end

precision is one of :native (default), :fp32, :bf16x3, :fp16x2, :ozaki_int8;
complex_mult is 3 (Gauss, default) or 4.  The backend methods also take them
as keyword arguments.  The extension assumes the definitions shown in
INTEGRATION.md next to the other backends (src/Backend.jl:6-14) and domains
(src/Domain.jl:6-9).  Set ENV["MUSCLE_HIP_LIB"] to the library path.
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

# Library handles, pooled per device.  A call checks a handle out for its
# duration, so no two tasks use one handle at once (a handle's stream is set
# before every call and its workspace is shared by its calls), and checks it
# back in; handles outlive the tasks that used them and are destroyed at exit.
# The stream is AMDGPU's task-local stream; the library orders a handle's new
# stream after the one it last used.
const HANDLES = Dict{Int,Vector{Ptr{Cvoid}}}()
const HANDLES_LOCK = ReentrantLock()

function with_handle(f)
    dev = AMDGPU.device_id(AMDGPU.device())
    h = lock(HANDLES_LOCK) do
        free = get!(HANDLES, dev, Ptr{Cvoid}[])
        isempty(free) || return pop!(free)
        ref = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:muscle_hip_create, LIB), Cint, (Ref{Ptr{Cvoid}},), ref))
        return ref[]
    end
    try
        check(ccall((:muscle_hip_set_stream, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), h, AMDGPU.stream().stream))
        return f(h)
    finally
        lock(() -> push!(HANDLES[dev], h), HANDLES_LOCK)
    end
end

function __init__()
    atexit() do
        lock(HANDLES_LOCK) do
            for hs in values(HANDLES), h in hs
                ccall((:muscle_hip_destroy, LIB), Cint, (Ptr{Cvoid},), h)
            end
            empty!(HANDLES)
        end
    end
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

const PRECISION = Dict{Symbol,Cint}(:native => 0, :fp32 => 1, :bf16x3 => 2, :fp16x2 => 3, :ozaki_int8 => 4)

function contract!(c::Tensor, a::Tensor, b::Tensor; precision::Symbol=:native, target_rel_error::Real=1e-12,
                   complex_mult::Integer=3)
    prec = get(PRECISION, precision, nothing)
    isnothing(prec) && throw(ArgumentError("unknown precision $precision"))
    complex_mult in (3, 4) || throw(ArgumentError("complex_mult must be 3 or 4"))
    labels = modemap(a, b, c)
    dc, kc = describe(c, labels)
    da, ka = describe(a, labels)
    db, kb = describe(b, labels)
    with_handle() do h
        check(ccall((:muscle_hip_set_option, LIB), Cint, (Ptr{Cvoid}, Cint, Int64), h, 0, complex_mult))
        GC.@preserve c a b kc ka kb begin
            check(ccall((:muscle_hip_binary_einsum, LIB), Cint,
                        (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}, Cint, Cdouble, Ptr{CInfo}),
                        h, dc, da, db, prec, target_rel_error, C_NULL))
        end
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

function Muscle.binary_einsum(be::BackendHIP, inds_c, a::Tensor, b::Tensor; precision=be.precision,
                              target_rel_error=be.target_rel_error, complex_mult=be.complex_mult)
    sizes = Dict{Index,Int}()
    for t in (a, b), (i, n) in zip(inds(t), size(parent(t)))
        sizes[i] = n
    end
    T = Base.promote_eltype(parent(a), parent(b))
    c = Tensor(similar(parent(a), T, Tuple(sizes[i] for i in inds_c)), collect(inds_c))
    return contract!(c, a, b; precision, target_rel_error, complex_mult)
end

Muscle.binary_einsum!(be::BackendHIP, c::Tensor, a::Tensor, b::Tensor; precision=be.precision,
                      target_rel_error=be.target_rel_error, complex_mult=be.complex_mult) =
    contract!(c, a, b; precision, target_rel_error, complex_mult)

function Muscle.hadamard!(::BackendHIP, c::Tensor, a::Tensor, b::Tensor)
    labels = modemap(a, b)
    dc, kc = describe(c, labels)
    da, ka = describe(a, labels)
    db, kb = describe(b, labels)
    with_handle() do h
        GC.@preserve c a b kc ka kb begin
            check(ccall((:muscle_hip_hadamard, LIB), Cint,
                        (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}, Ptr{CInfo}), h, dc, da, db, C_NULL))
        end
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
    with_handle() do h
        GC.@preserve y x ky kx begin
            check(ccall((:muscle_hip_unary_einsum, LIB), Cint, (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ptr{CInfo}),
                        h, dy, dx, C_NULL))
        end
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
Muscle.choose_backend_rule(::typeof(Muscle.tensor_svd_thin!), ::Vararg{DomainROCm,4}) = BackendHIP()
Muscle.choose_backend_rule(::typeof(Muscle.tensor_svd_trunc), ::DomainROCm) = BackendHIP()
Muscle.choose_backend_rule(::typeof(Muscle.tensor_svd_trunc!), ::Vararg{DomainROCm,4}) = BackendHIP()
Muscle.choose_backend_rule(::typeof(Muscle.tensor_qr_thin), ::DomainROCm) = BackendHIP()
Muscle.choose_backend_rule(::typeof(Muscle.tensor_qr_thin!), ::DomainROCm, ::DomainROCm, ::DomainROCm) = BackendHIP()
# simple_update's generic method (src/Operations/simple_update.jl:35-82) then
# runs its contractions, SVD and absorption through this backend
Muscle.choose_backend_rule(::typeof(Muscle.simple_update), ::DomainROCm, ::DomainROCm, ::DomainROCm) = BackendHIP()

# muscle_hip_svd_thin into caller-owned U[inds_u..., ind_s], s[ind_s], Vt[inds_v..., ind_s]
function svd_into!(U::Tensor, s::Tensor, Vt::Tensor, A::Tensor)
    labels = modemap(U, s, Vt, A)
    du, ku = describe(U, labels)
    ds, ks = describe(s, labels)
    dv, kv = describe(Vt, labels)
    da, ka = describe(A, labels)
    with_handle() do h
        GC.@preserve U s Vt A ku ks kv ka begin
            check(ccall((:muscle_hip_svd_thin, LIB), Cint,
                        (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}), h, du, ds, dv, da))
        end
    end
    return U, s, Vt
end

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
    return svd_into!(U, s, Vt, A)
end

# In place (src/Operations/tensor_svd.jl:129): the factors are written straight
# into the caller's U, s, V (the bond is s's index; U's other indices are the
# row indices), with no intermediate copy.
function Muscle.tensor_svd_thin!(::BackendHIP, U::Tensor, s::Tensor, V::Tensor, A::Tensor; kwargs...)
    ind_s = only(inds(s))
    inds_u, inds_v = setdiff(inds(U), [ind_s]), setdiff(inds(V), [ind_s])
    issetequal(inds_u ∪ inds_v, inds(A)) || throw(ArgumentError("inds(U) and inds(V) must cover inds(A)"))
    size(s, ind_s) == min(prod(i -> size(A, i), inds_u), prod(i -> size(A, i), inds_v)) ||
        throw(DimensionMismatch("s must hold min(rows, columns) singular values"))
    return svd_into!(U, s, V, A)
end

# the cut of src/Operations/tensor_svd.jl:183-203, on the host: at most maxdim
# values, then up to the first one below threshold * norm(s)
function trunc_dim(s::Tensor, threshold, maxdim)
    sh = Array(parent(s))
    k = length(sh)
    isnothing(maxdim) || (k = min(k, maxdim))
    isnothing(threshold) || (k = something(findfirst(<(sqrt(sum(abs2, sh)) * threshold), view(sh, 1:k)), k))
    return k
end

function Muscle.tensor_svd_trunc(::BackendHIP, A::Tensor; inds_u=(), inds_v=(), ind_s=Index(gensym(:vind)),
                                 threshold=nothing, maxdim=nothing, kwargs...)
    U, s, Vt = Muscle.tensor_svd_thin(BackendHIP(), A; inds_u, inds_v, ind_s)
    k = trunc_dim(s, threshold, maxdim)
    return view(U, ind_s => 1:k), view(s, ind_s => 1:k), view(Vt, ind_s => 1:k)
end

# In place (src/Operations/tensor_svd.jl:217): as BackendBase does, the
# truncated factors must have the caller's sizes and are copied into them.
function Muscle.tensor_svd_trunc!(::BackendHIP, U::Tensor, s::Tensor, V::Tensor, A::Tensor; threshold=nothing,
                                  maxdim=nothing, kwargs...)
    ind_s = only(inds(s))
    tU, ts, tV = Muscle.tensor_svd_trunc(BackendHIP(), A; inds_u=setdiff(inds(U), [ind_s]),
                                         inds_v=setdiff(inds(V), [ind_s]), ind_s, threshold, maxdim)
    inds(tU) == inds(U) && inds(ts) == inds(s) && inds(tV) == inds(V) ||
        throw(ArgumentError("output indices differ from the factorization's"))
    size(tU) == size(U) && size(ts) == size(s) && size(tV) == size(V) ||
        throw(DimensionMismatch("truncated factors have sizes $(size(tU)), $(size(ts)), $(size(tV))"))
    copyto!(parent(U), parent(tU))
    copyto!(parent(s), parent(ts))
    copyto!(parent(V), parent(tV))
    return U, s, V
end

# muscle_hip_qr_thin into caller-owned Q[inds_q..., v], R[v, inds_r...]
function qr_into!(Q::Tensor, R::Tensor, A::Tensor)
    labels = modemap(Q, R, A)
    dq, kq = describe(Q, labels)
    dr, kr = describe(R, labels)
    da, ka = describe(A, labels)
    with_handle() do h
        GC.@preserve Q R A kq kr ka begin
            check(ccall((:muscle_hip_qr_thin, LIB), Cint, (Ptr{Cvoid}, Ref{CTensor}, Ref{CTensor}, Ref{CTensor}),
                        h, dq, dr, da))
        end
    end
    return Q, R
end

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
    return qr_into!(Q, R, A)
end

# In place (src/Operations/tensor_qr.jl:70): written straight into the caller's
# Q and R; the bond is the index they share.
function Muscle.tensor_qr_thin!(::BackendHIP, Q::Tensor, R::Tensor, A::Tensor; kwargs...)
    bond = only(intersect(inds(Q), inds(R)))
    issetequal(setdiff(inds(Q), [bond]) ∪ setdiff(inds(R), [bond]), inds(A)) ||
        throw(ArgumentError("inds(Q) and inds(R) must cover inds(A)"))
    return qr_into!(Q, R, A)
end

end # module
