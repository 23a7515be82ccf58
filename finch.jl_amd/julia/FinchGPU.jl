# FinchGPU.jl — Julia side of the MI355X backend: the binding Finch.jl would
# load to reach libfinchgpu.so.  It mirrors finchgpu.h one to one and plugs
# into Finch at the three points the backend replaces:
#
#   * a device, `GPU <: Finch.AbstractDevice` (src/architecture.jl:6-64,170-173),
#     with `transfer` of level buffers (src/architecture.jl:163,269-294);
#   * device-resident level buffers, `GPUVector{T} <: AbstractVector{T}`, used
#     as the Ptr/Idx/Val parameters of SparseList/Element levels
#     (src/tensors/levels/sparse_list_levels.jl:37-42, element_levels.jl:20-22);
#   * routing of the recognised programs before code generation: the
#     `spmv!`/`spmm!`/`sddmm!`/`ewise`/`spgemm` methods below are what the
#     staged `execute` (src/execute.jl:101-126) and `LogicExecutor`
#     (src/scheduler/LogicExecutor.jl:86-96) call when a program's tensors
#     carry `GPUVector` buffers.
#
# Julia is not part of this build's toolchain, so this file is the reference
# binding, exercised only where Julia is installed; the same C ABI is
# exercised from Python (finchgpu/_lib.py) by the tests.
module FinchGPU

using Finch
using Finch: AbstractDevice, DenseLevel, SparseListLevel, ElementLevel, Tensor, SwizzleArray

const libfgx = joinpath(@__DIR__, "..", "finchgpu", "libfinchgpu.so")
const libhip = "libamdhip64"

# ---- C structs (finchgpu.h) -------------------------------------------------------
struct FgxCompressed
    n_outer::Int64
    n_inner::Int64
    nnz::Int64
    ptr_type::Int32
    idx_type::Int32
    val_type::Int32
    index_base::Int32
    ptr::Ptr{Cvoid}
    idx::Ptr{Cvoid}
    val::Ptr{Cvoid}
end

const FGX_OK, FGX_EDIM, FGX_EPROTO, FGX_EUNSUP, FGX_EOOM, FGX_EADMIT, FGX_EHIP, FGX_EINVAL = 0:7
const FGX_F64, FGX_F32, FGX_BF16, FGX_F16, FGX_I32, FGX_I64 = Int32.(0:5)
const FGX_ROW_MAJOR, FGX_COL_MAJOR = Int32(0), Int32(1)

fgx_dtype(::Type{Float64}) = FGX_F64
fgx_dtype(::Type{Float32}) = FGX_F32
fgx_dtype(::Type{Int32}) = FGX_I32
fgx_dtype(::Type{Int64}) = FGX_I64

struct FinchGPUError <: Exception
    code::Int
    msg::String
end

# fgx_status -> the exception Finch itself raises for the same condition.
function check(code)
    code == FGX_OK && return nothing
    msg = unsafe_string(ccall((:fgx_last_error, libfgx), Cstring, ()))
    code == FGX_EDIM && throw(DimensionMismatch(msg))
    code == FGX_EPROTO && throw(Finch.FinchProtocolError(msg))
    throw(FinchGPUError(code, msg))
end

# ---- device and buffers ---------------------------------------------------------------
struct GPU <: AbstractDevice
    id::Int
    ctx::Ptr{Cvoid}
end

function GPU(id::Integer = 0)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fgx_ctx_create, libfgx), Cint, (Cint, Ref{Ptr{Cvoid}}), id, ctx))
    GPU(id, ctx[])
end

Finch.get_num_tasks(::GPU) = 1

"""Release the context (its stream, scratch and plan cache); `dev` is unusable afterwards."""
close!(dev::GPU) = check(ccall((:fgx_ctx_destroy, libfgx), Cint, (Ptr{Cvoid},), dev.ctx))

"""The context's own non-blocking stream (a hipStream_t), for callers that want one."""
stream(dev::GPU) = ccall((:fgx_ctx_stream, libfgx), Ptr{Cvoid}, (Ptr{Cvoid},), dev.ctx)

"""Wait for the work queued on `stream` (C_NULL: the HIP null stream)."""
sync(dev::GPU, stream::Ptr{Cvoid} = C_NULL) =
    check(ccall((:fgx_sync, libfgx), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), dev.ctx, stream))

"""ABI version of the loaded library (`FGX_ABI_VERSION` in include/finchgpu.h)."""
abi_version() = ccall((:fgx_version, libfgx), Cint, ())

"""Drop the context's cached per-operand plans (tile tables): needed only after
an operand's `ptr` was freed and its address reused, to release memory early."""
clear_plans!(dev::GPU) = check(ccall((:fgx_plan_cache_clear, libfgx), Cint, (Ptr{Cvoid},), dev.ctx))

"""Check a Dense(SparseList(Element)) operand against the SparseList invariants
on the device (fgx_validate): `FinchProtocolError` for a malformed `ptr` or an
unsorted fibre, `DimensionMismatch` for coordinates outside the shape."""
validate(dev::GPU, A) = check(ccall((:fgx_validate, libfgx), Cint,
                                    (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}),
                                    dev.ctx, descriptor(A), C_NULL))

"""Finch's `mode=:safe` for every later call through `dev`: each compressed
operand is validated before a kernel reads it (`safe=false`: trusted, as `:fast`)."""
set_safe!(dev::GPU, safe::Bool) =
    check(ccall((:fgx_ctx_set_safe, libfgx), Cint, (Ptr{Cvoid}, Cint), dev.ctx, safe))

"""A device vector: the storage of a level buffer on the GPU (hipMalloc'd)."""
mutable struct GPUVector{T} <: AbstractVector{T}
    ptr::Ptr{T}
    len::Int
    function GPUVector{T}(len::Integer) where {T}
        p = Ref{Ptr{Cvoid}}(C_NULL)
        ccall((:hipMalloc, libhip), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, max(len, 1) * sizeof(T)) == 0 ||
            throw(OutOfMemoryError())
        v = new{T}(Ptr{T}(p[]), len)
        finalizer(x -> ccall((:hipFree, libhip), Cint, (Ptr{Cvoid},), x.ptr), v)
        v
    end
end
Base.size(v::GPUVector) = (v.len,)
Base.getindex(::GPUVector, ::Int) = error("GPUVector: element access on the host is not supported")
Base.pointer(v::GPUVector) = v.ptr
function Base.resize!(v::GPUVector{T}, n::Integer) where {T}
    n <= v.len && (v.len = n; return v)
    w = GPUVector{T}(n)
    ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint), w.ptr, v.ptr, v.len * sizeof(T), 3)
    v.ptr, v.len = w.ptr, n
    return v
end

function Finch.transfer(::GPU, v::Vector{T}) where {T}
    w = GPUVector{T}(length(v))
    ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint), w.ptr, v, length(v) * sizeof(T), 1)
    w
end

# ---- operand descriptors ----------------------------------------------------------------
# A Dense(SparseList(Element)) tensor with GPUVector buffers; swizzle(A, 2, 1)
# is the same descriptor read row-wise.
function descriptor(A::Tensor)
    dense = A.lvl
    sl = dense.lvl
    el = sl.lvl
    FgxCompressed(dense.shape, sl.shape, length(sl.idx), fgx_dtype(eltype(sl.ptr)),
                  fgx_dtype(eltype(sl.idx)), fgx_dtype(eltype(el.val)), 1,
                  pointer(sl.ptr), pointer(sl.idx), pointer(el.val))
end

# ---- routed programs ------------------------------------------------------------------------
# y .= 0; for j = _, i = _; y[i] += A[i, j] * x[j]   (test/reference64/typical/typical_spmv_csc.txt)
function spmv!(dev::GPU, y::GPUVector, A::Tensor, x::GPUVector)
    d = descriptor(A)
    check(ccall((:fgx_spmv_csc, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, d, x.ptr, y.ptr, C_NULL))
    y
end

# y .= 0; for i = _, j = _; y[i] += A[j, i] * x[j] over swizzle(A, 2, 1)
function spmv!(dev::GPU, y::GPUVector, S::SwizzleArray{(2, 1)}, x::GPUVector)
    d = descriptor(S.body)
    check(ccall((:fgx_spmv_csr, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, d, x.ptr, y.ptr, C_NULL))
    y
end

# The same CSR SpMV through the entry point that allocates nothing, for code
# under HIP graph capture or loops that may not allocate; its scratch is a
# fixed per-GPU array of FGX_NOALLOC_SCRATCH_BYTES, shared by every call.
function spmv_noalloc!(dev::GPU, y::GPUVector, S::SwizzleArray{(2, 1)}, x::GPUVector;
                       stream::Ptr{Cvoid} = C_NULL)
    d = descriptor(S.body)
    check(ccall((:fgx_spmv_csr_noalloc, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, d, x.ptr, y.ptr, stream))
    y
end

# @einsum C[i, k] += A[i, j] * B[j, k]; B, C column-major (Finch's Dense(Dense)).
function spmm!(dev::GPU, C::GPUVector, S::SwizzleArray{(2, 1)}, B::GPUVector, k::Integer)
    d = descriptor(S.body)
    check(ccall((:fgx_spmm_csr, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Int64, Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}, Int64,
                 Cint, Ptr{Cvoid}),
                dev.ctx, d, k, B.ptr, d.n_inner, FGX_COL_MAJOR, C.ptr, d.n_outer, FGX_COL_MAJOR, C_NULL))
    C
end

# compute(lazy(M) .* (lazy(X) * lazy(Y))): X m x k, Y k x n (column-major,
# i.e. Y^T row-major with leading dimension k).
function sddmm!(dev::GPU, out::GPUVector, S::SwizzleArray{(2, 1)}, X::GPUVector, Y::GPUVector, k::Integer)
    d = descriptor(S.body)
    check(ccall((:fgx_sddmm_csr, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Int64, Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}, Int64, Cint,
                 Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, d, k, X.ptr, d.n_outer, FGX_COL_MAJOR, Y.ptr, k, FGX_ROW_MAJOR, out.ptr, C_NULL))
    out
end

# Two-phase sparse outputs: symbolic -> resize! (freeze) -> numeric, as the
# SparseList level's assemble/freeze protocol (sparse_list_levels.jl:301-327).
function ewise(dev::GPU, op::Symbol, A::Tensor, B::Tensor)
    opc = Dict(:* => 0, :+ => 1, :max => 2)[op]
    da, db = descriptor(A), descriptor(B)
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(da.n_outer + 1)
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_ewise_symbolic, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ref{Int64}, Ptr{Cvoid}),
                dev.ctx, opc, da, db, ptr.ptr, nnz, C_NULL))
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(nnz[])
    val = GPUVector{eltype(A.lvl.lvl.lvl.val)}(nnz[])
    check(ccall((:fgx_ewise_numeric, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}),
                dev.ctx, opc, da, db, ptr.ptr, idx.ptr, val.ptr, C_NULL))
    Tensor(DenseLevel(SparseListLevel(ElementLevel{zero(eltype(val))}(val), A.lvl.lvl.shape, ptr, idx),
                      A.lvl.shape))
end

# Gustavson C = A * B (column-compressed operands: C = (B^T A^T)^T, the
# reference's column order, docs/examples/spgemm.jl:59-76).
function spgemm(dev::GPU, A::Tensor, B::Tensor; semiring::Symbol = :plus_times)
    sr = semiring === :min_plus ? 1 : 0
    db, da = descriptor(B), descriptor(A)
    ptr = GPUVector{eltype(B.lvl.lvl.ptr)}(db.n_outer + 1)
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_spgemm_symbolic, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ref{Int64}, Ptr{Cvoid}),
                dev.ctx, sr, db, da, ptr.ptr, nnz, C_NULL))
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(nnz[])
    val = GPUVector{eltype(A.lvl.lvl.lvl.val)}(nnz[])
    check(ccall((:fgx_spgemm_numeric, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}),
                dev.ctx, sr, db, da, ptr.ptr, idx.ptr, val.ptr, C_NULL))
    fill = semiring === :min_plus ? typemax(eltype(val)) : zero(eltype(val))
    Tensor(DenseLevel(SparseListLevel(ElementLevel{fill}(val), A.lvl.lvl.shape, ptr, idx), B.lvl.shape))
end

# ---- format conversion on the device --------------------------------------------------------
# permutedims(A) of a Dense(SparseList) tensor (src/interface/copy.jl:64-71):
# the same matrix stored along its other dimension; the transpose concordize
# inserts (src/scheduler/optimize.jl:396-452) when loop and storage order differ.
function Base.permutedims(dev::GPU, A::Tensor)
    d = descriptor(A)
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(d.n_inner + 1)
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(d.nnz)
    val = GPUVector{eltype(A.lvl.lvl.lvl.val)}(d.nnz)
    check(ccall((:fgx_transpose, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, d, ptr.ptr, idx.ptr, val.ptr, C_NULL))
    Tensor(DenseLevel(SparseListLevel(ElementLevel{Finch.fill_value(A)}(val), A.lvl.shape, ptr, idx),
                      A.lvl.lvl.shape))
end

# Tensor(Dense(SparseList(Element(0))), fsparse(I, J, V, (m, n))) assembled on
# the device (src/interface/fsparse.jl:64-97): column fibres, rows sorted,
# duplicates summed left to right in input order.
function fsparse_csc(dev::GPU, I::GPUVector{Ti}, J::GPUVector{Ti}, V::GPUVector{Tv},
                     shape::Tuple{Integer,Integer}) where {Ti,Tv}
    m, n = shape
    ptr = GPUVector{Ti}(n + 1)
    idx = GPUVector{Ti}(length(V))
    val = GPUVector{Tv}(length(V))
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_coo_to_compressed, libfgx), Cint,
                (Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Int64}, Ptr{Cvoid}),
                dev.ctx, n, m, length(V), fgx_dtype(Ti), fgx_dtype(Tv), 1, J.ptr, I.ptr, V.ptr,
                ptr.ptr, idx.ptr, val.ptr, nnz, C_NULL))
    resize!(idx, nnz[]); resize!(val, nnz[])       # freeze: shrink to the distinct count
    Tensor(DenseLevel(SparseListLevel(ElementLevel{zero(Tv)}(val), m, ptr, idx), n))
end

# dropfills(A) (src/interface/copy.jl:73-117) into the same format.
function Finch.dropfills(dev::GPU, A::Tensor)
    d = descriptor(A)
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(d.n_outer + 1)
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(d.nnz)
    val = GPUVector{eltype(A.lvl.lvl.lvl.val)}(d.nnz)
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_dropfills, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ref{Int64}, Ptr{Cvoid}),
                dev.ctx, d, Float64(Finch.fill_value(A)), ptr.ptr, idx.ptr, val.ptr, nnz, C_NULL))
    resize!(idx, nnz[]); resize!(val, nnz[])
    Tensor(DenseLevel(SparseListLevel(ElementLevel{Finch.fill_value(A)}(val), A.lvl.lvl.shape, ptr, idx),
                      A.lvl.shape))
end

# ---- value-precision form -------------------------------------------------------------
# The BF16/F16 storage form of an operand's values (FP32 accumulation in every
# kernel).  Finch has no 16-bit float element in this path; the converted
# buffer backs an ElementLevel whose values the kernels widen on load.
function convert_values!(dev::GPU, dst::GPUVector, dst_type::Int32, src::GPUVector, src_type::Int32)
    check(ccall((:fgx_convert, libfgx), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}),
                dev.ctx, src.ptr, src_type, dst.ptr, dst_type, length(src), C_NULL))
    dst
end

# ---- batched form ---------------------------------------------------------------------
struct FgxBatchItem           # struct fgx_batch_item
    A::FgxCompressed
    n_rhs::Int64
    B::Ptr{Cvoid}
    C::Ptr{Cvoid}
end

const FGX_BATCH_FORCE_ROWSPLIT, FGX_BATCH_FORCE_NNZSPLIT, FGX_BATCH_NO_GRAPH = 1, 2, 4

"""Many independent `C = A*B` (CSR A, row-major F32 B/C with `n_rhs` columns)
in one persistent launch, replayed from a hipGraph.  The items' buffers must
outlive the plan."""
mutable struct BatchPlan
    dev::GPU
    handle::Ptr{Cvoid}
    items::Vector{FgxBatchItem}
end

function BatchPlan(dev::GPU, items::Vector{FgxBatchItem}; flags::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:fgx_batch_plan_create, libfgx), Cint,
                (Ptr{Cvoid}, Ptr{FgxBatchItem}, Int64, Cint, Ref{Ptr{Cvoid}}),
                dev.ctx, items, length(items), flags, h))
    plan = BatchPlan(dev, h[], items)
    finalizer(p -> ccall((:fgx_batch_plan_destroy, libfgx), Cint, (Ptr{Cvoid},), p.handle), plan)
    plan
end

run!(plan::BatchPlan) = check(ccall((:fgx_batch_plan_run, libfgx), Cint,
                                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                                    plan.dev.ctx, plan.handle, C_NULL))

"""[work items, row-split items, nnz-split items, grid, nnz, algorithmic bytes per run]."""
function stats(plan::BatchPlan)
    st = zeros(Int64, 6)
    check(ccall((:fgx_batch_plan_stats, libfgx), Cint, (Ptr{Cvoid}, Ptr{Int64}), plan.handle, st))
    st
end

"""The plain path: one SpMM (SpMV for n_rhs == 1) call per item."""
run_loop!(dev::GPU, items::Vector{FgxBatchItem}) =
    check(ccall((:fgx_batch_run_loop, libfgx), Cint,
                (Ptr{Cvoid}, Ptr{FgxBatchItem}, Int64, Ptr{Cvoid}),
                dev.ctx, items, length(items), C_NULL))

# ---- serving loop ---------------------------------------------------------------------
const FGX_REQ_SPMV, FGX_REQ_SPMM, FGX_REQ_SDDMM = Int32(0), Int32(1), Int32(2)

struct FgxRequest             # struct fgx_request
    kind::Int32
    operand::Int32
    k::Int64
    arrival_s::Float64
    in::Ptr{Cvoid}
    in2::Ptr{Cvoid}
    out::Ptr{Cvoid}
end

struct FgxServeConfig         # struct fgx_serve_config
    hbm_budget_bytes::Int64
    max_slots::Int32
    max_batch::Int32
    mode::Int32               # 0: continuous batching, 1: plain per-call path
    reserved::Int32
end

struct FgxRequestResult       # struct fgx_request_result
    admit_s::Float64
    done_s::Float64
    status::Int32
    batch::Int32
end

"""Replay an open-loop request trace against device-resident operands
(continuous batching with HBM admission); returns one result per request."""
function serve(dev::GPU, operands::Vector{FgxCompressed}, reqs::Vector{FgxRequest};
               hbm_budget_bytes::Integer = 200 * 10^9, max_slots::Integer = 4096,
               max_batch::Integer = 512, plain::Bool = false)
    cfg = FgxServeConfig(hbm_budget_bytes, max_slots, max_batch, plain ? 1 : 0, 0)
    res = Vector{FgxRequestResult}(undef, length(reqs))
    check(ccall((:fgx_serve_run, libfgx), Cint,
                (Ptr{Cvoid}, Ptr{FgxCompressed}, Int64, Ptr{FgxRequest}, Int64,
                 Ref{FgxServeConfig}, Ptr{FgxRequestResult}),
                dev.ctx, operands, length(operands), reqs, length(reqs), cfg, res))
    res
end

end # module
