# FinchGPU.jl — Julia side of the MI355X backend: the package a Finch.jl user
# loads to run Finch's sparse hot loops on libfinchgpu.so.  It plugs into
# Finch at the points the backend replaces:
#
#   * a device, `GPU <: Finch.AbstractDevice` (src/architecture.jl:6-64,170-173),
#     with its memory spaces (`local_memory`, `shared_memory`,
#     `global_memory`, src/architecture.jl:9-28,213-250) and `transfer` of
#     level buffers (src/architecture.jl:163,269-294);
#   * device-resident level buffers, `GPUVector{T} <: AbstractVector{T}`, used
#     as the Ptr/Idx/Val parameters of SparseList/SparseDict/Element levels
#     (src/tensors/levels/sparse_list_levels.jl:37-42, element_levels.jl:20-22);
#   * the `@finch` path: a method of `Finch.execute` for program instances
#     (src/execute.jl:101-126 stages `execute_impl`; the `@finch` macro calls
#     `execute`, src/execute.jl:180-220) that sends a program over GPU-resident
#     tensors to the C entry point its routing-table entry names;
#   * the `compute` path: `GPUScheduler`, installed with `Finch.set_scheduler!`
#     (src/interface/lazy.jl:666-696) in place of the `LogicExecutor`
#     (src/scheduler/LogicExecutor.jl:86-96), which matches the logic program
#     of `compute`, `@einsum`, `A * B`, `A .* B`, `max.(A, B)` against the same
#     table and returns `Dense(SparseDict)` results as Finch's scheduler does
#     (src/interface/traits.jl:509-517).
#
# Both routes read ../routes.toml, the table the Python mirror
# (finchgpu/interface.py) reads too; a program over GPU-resident operands with
# no route raises FinchGPUError(FGX_EUNSUP) and never falls back to the host.
module FinchGPU

using Finch
using TOML
using Finch: AbstractDevice, Tensor, SwizzleArray, DenseLevel, SparseListLevel, SparseDictLevel,
    ElementLevel, AtomicElementLevel, MutexLevel, SeparateLevel
using Finch.FinchNotation: FinchNodeInstance, BlockInstance, LoopInstance, AssignInstance,
    AccessInstance, CallInstance, TagInstance, LiteralInstance, IndexInstance, DeclareInstance,
    FreezeInstance, ThawInstance, YieldBindInstance, VariableInstance, UpdaterInstance
using Finch.FinchLogic: LogicNode, plan, query, produces, alias, table, immediate, field,
    mapjoin, aggregate, reorder, relabel, reformat, subquery

const libfgx = joinpath(@__DIR__, "..", "finchgpu", "libfinchgpu.so")
const libhip = "libamdhip64"
const ROUTES_FILE = joinpath(@__DIR__, "..", "routes.toml")

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
fgx_dtype(::Type{Float16}) = FGX_F16
fgx_dtype(::Type{Int32}) = FGX_I32
fgx_dtype(::Type{Int64}) = FGX_I64
fgx_dtype(::Type{T}) where {T} = throw(FinchGPUError(FGX_EUNSUP, "element type $T"))

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

unsupported(msg) = throw(FinchGPUError(FGX_EUNSUP, msg))

# ---- device and memory spaces ---------------------------------------------------------
"""
    GPU(id = 0)

One HIP device (one process per GPU): Finch's `AbstractDevice` for the
backend.  Holds the library context (its streams' scratch, the tile-table
cache); one instance per device id, shared by every tensor on that device.
"""
struct GPU <: AbstractDevice
    id::Int
    ctx::Ptr{Cvoid}
end

const DEVICES = Dict{Int,GPU}()
const DEVICES_LOCK = ReentrantLock()

function GPU(id::Integer = 0)
    lock(DEVICES_LOCK) do
        get!(DEVICES, Int(id)) do
            ctx = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:fgx_ctx_create, libfgx), Cint, (Cint, Ref{Ptr{Cvoid}}), id, ctx))
            GPU(Int(id), ctx[])
        end
    end
end

# A launch spreads over the whole device; Finch's task count is the number of
# host tasks that drive it, one.
Finch.get_num_tasks(::GPU) = 1

"""HBM of one GPU: where every level buffer of a GPU-resident tensor lives.
Per-workgroup memory (LDS) exists only inside the kernels, so the local and
shared spaces of a `GPU` are this same global space on the host side."""
struct GPUGlobalMemory
    device::GPU
end
Finch.local_memory(dev::GPU) = GPUGlobalMemory(dev)
Finch.shared_memory(dev::GPU) = GPUGlobalMemory(dev)
Finch.global_memory(dev::GPU) = GPUGlobalMemory(dev)

# Parallel regions.  Finch lowers `for i = parallel(_, device)` through
# virtual_parallel_region (src/architecture.jl:685-824), which for a CPU
# spreads the loop over host tasks.  A program over GPU-resident tensors never
# reaches lowering: the execute hook below sends the whole program to its
# routed kernel, whose launch covers the device.  A GPU handed to a parallel
# loop of a program that is not routed therefore stops at lowering with
# FGX_EUNSUP instead of generating host code for a device it cannot drive.
struct VirtualGPU <: Finch.AbstractVirtualDevice end
Finch.virtualize(ctx, ex, ::Type{GPU}) = VirtualGPU()
Finch.virtual_parallel_region(f, ctx, ext, ::VirtualGPU, schedule) =
    unsupported("a parallel region on a GPU device: only routed programs run on the GPU")

"""Release the context (its scratch and plan cache); `dev` is unusable afterwards."""
function close!(dev::GPU)
    lock(DEVICES_LOCK) do
        delete!(DEVICES, dev.id)
    end
    check(ccall((:fgx_ctx_destroy, libfgx), Cint, (Ptr{Cvoid},), dev.ctx))
end

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

"""Finch's `mode=:safe` for every later call through `dev`: each compressed
operand is validated before a kernel reads it (`safe=false`: trusted, as `:fast`)."""
set_safe!(dev::GPU, safe::Bool) =
    check(ccall((:fgx_ctx_set_safe, libfgx), Cint, (Ptr{Cvoid}, Cint), dev.ctx, safe))

# ---- device buffers -----------------------------------------------------------------
function hip_alloc(bytes::Integer)
    p = Ref{Ptr{Cvoid}}(C_NULL)
    ccall((:hipMalloc, libhip), Cint, (Ref{Ptr{Cvoid}}, Csize_t), p, max(bytes, 1)) == 0 ||
        throw(OutOfMemoryError())
    p[]
end
hip_free(p::Ptr) = p == C_NULL || ccall((:hipFree, libhip), Cint, (Ptr{Cvoid},), p)
function hip_copy(dst::Ptr, src::Ptr, bytes::Integer, kind::Integer)
    bytes == 0 && return nothing
    ccall((:hipMemcpy, libhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Csize_t, Cint),
          dst, src, bytes, kind) == 0 || throw(FinchGPUError(FGX_EHIP, "hipMemcpy failed"))
    nothing
end
const H2D, D2H, D2D = 1, 2, 3

"""
    GPUVector{T}(dev, n)

A level buffer in HBM.  It owns exactly one hipMalloc'd block (`cap`
elements), freed once by its finalizer; `resize!` grows it by moving the
contents into a new block and freeing the old one, so no two objects ever
hold the same block.
"""
mutable struct GPUVector{T} <: AbstractVector{T}
    ptr::Ptr{T}
    len::Int
    cap::Int
    dev::GPU
    function GPUVector{T}(dev::GPU, len::Integer) where {T}
        v = new{T}(Ptr{T}(hip_alloc(len * sizeof(T))), len, len, dev)
        finalizer(release!, v)
        v
    end
end
GPUVector{T}(len::Integer) where {T} = GPUVector{T}(GPU(0), len)

function release!(v::GPUVector)
    p = v.ptr
    v.ptr, v.len, v.cap = Ptr{eltype(v)}(C_NULL), 0, 0
    hip_free(p)
    nothing
end

Base.size(v::GPUVector) = (v.len,)
Base.pointer(v::GPUVector) = v.ptr
Base.IndexStyle(::Type{<:GPUVector}) = IndexLinear()
# Element access from the host copies one element; the routed programs never
# use it, but Finch's printing and `countstored` read single entries of ptr.
function Base.getindex(v::GPUVector{T}, i::Int) where {T}
    @boundscheck checkbounds(v, i)
    r = Ref{T}()
    hip_copy(Base.unsafe_convert(Ptr{T}, r), v.ptr + (i - 1) * sizeof(T), sizeof(T), D2H)
    r[]
end
function Base.setindex!(v::GPUVector{T}, x, i::Int) where {T}
    @boundscheck checkbounds(v, i)
    r = Ref{T}(convert(T, x))
    hip_copy(v.ptr + (i - 1) * sizeof(T), Base.unsafe_convert(Ptr{T}, r), sizeof(T), H2D)
    v
end
function Base.resize!(v::GPUVector{T}, n::Integer) where {T}
    n < 0 && throw(ArgumentError("new length must be ≥ 0"))
    if n <= v.cap
        v.len = n
        return v
    end
    p = Ptr{T}(hip_alloc(n * sizeof(T)))
    hip_copy(p, v.ptr, v.len * sizeof(T), D2D)
    old = v.ptr
    v.ptr, v.len, v.cap = p, n, n
    hip_free(old)
    return v
end
function Base.Array(v::GPUVector{T}) where {T}
    a = Vector{T}(undef, v.len)
    hip_copy(pointer(a), v.ptr, v.len * sizeof(T), D2H)
    a
end
Base.similar(v::GPUVector, ::Type{T}, dims::Tuple{Int}) where {T} = GPUVector{T}(v.dev, dims[1])

function upload(dev::GPU, v::AbstractVector{T}) where {T}
    w = GPUVector{T}(dev, length(v))
    hv = collect(v)
    hip_copy(w.ptr, pointer(hv), length(hv) * sizeof(T), H2D)
    w
end

# transfer(device or memory, buffer): Finch's level transfers recurse down to
# here (sparse_list_levels.jl:64-69, sparse_dict_levels.jl:93-104,
# dense_levels.jl:47-49, element_levels.jl:46-48, tensors.jl:346).
Finch.transfer(dev::GPU, v::Vector) = upload(dev, v)
Finch.transfer(mem::GPUGlobalMemory, v::Vector) = upload(mem.device, v)
Finch.transfer(dev::GPU, v::GPUVector) = v.dev == dev ? v : upload(dev, Array(v))
Finch.transfer(mem::GPUGlobalMemory, v::GPUVector) = Finch.transfer(mem.device, v)
# SparseDict's insertion table and pool stay on the host.
Finch.transfer(::Union{GPU,GPUGlobalMemory}, d::Dict) = d
# The levels of Finch's threaded loops (atomic_element_levels.jl:15-48,
# mutex_levels.jl:19-41, separate_levels.jl:22-37).  AtomicElement's values
# move like Element's.  A Mutex's locks guard host threads; they stay on the
# host, since on the device every update of a routed output is a device
# atomic.  A Separate level moves each position's own level.
Finch.transfer(dev::Union{GPU,GPUGlobalMemory}, lvl::AtomicElementLevel{Vf,Tv,Tp}) where {Vf,Tv,Tp} =
    AtomicElementLevel{Vf,Tv,Tp}(Finch.transfer(dev, lvl.val))
Finch.transfer(::Union{GPU,GPUGlobalMemory}, locks::Vector{<:Base.AbstractLock}) = locks
Finch.transfer(dev::Union{GPU,GPUGlobalMemory}, lvls::Vector{<:Finch.AbstractLevel}) =
    [Finch.transfer(dev, l) for l in lvls]

"""The host: `transfer(Host(), tensor)` copies every level buffer back."""
struct Host end
Finch.transfer(::Host, v::GPUVector) = Array(v)
Finch.transfer(::Host, v::AbstractArray) = v
Finch.transfer(::Host, d::Dict) = d

on_gpu(v::GPUVector) = true
on_gpu(v) = false

# ---- operand descriptors ----------------------------------------------------------------
# Positions written by the kernels: a SparseDict built here from a device
# result is in coordinate order (val == 1:nnz); other SparseDict buffers are
# not handed to the kernels.
const IN_ORDER = WeakKeyDict{GPUVector,Bool}()

sparse_level(A::Tensor) = A.lvl isa DenseLevel && (A.lvl.lvl isa SparseListLevel ||
                                                   A.lvl.lvl isa SparseDictLevel) &&
                          A.lvl.lvl.lvl isa ElementLevel
# A value leaf: Element or AtomicElement, possibly under Mutex locks.  As the
# output of a routed program, AtomicElement and Mutex(Element) receive the
# kernel's result like Element: the device kernels' own update protocol
# (device atomics for the column scatter, one writer per row otherwise)
# stands in for the per-element atomics and locks.
value_leaf(l) = l isa ElementLevel || l isa AtomicElementLevel ||
                (l isa MutexLevel && value_leaf(l.lvl))
dense_level(A::Tensor) = (A.lvl isa DenseLevel && value_leaf(A.lvl.lvl)) ||
                         (A.lvl isa DenseLevel && A.lvl.lvl isa DenseLevel &&
                          value_leaf(A.lvl.lvl.lvl))
# Dense(Separate(Element)): each position owns its own Element level.
separate_level(A::Tensor) = A.lvl isa DenseLevel && A.lvl.lvl isa SeparateLevel &&
                            A.lvl.lvl.lvl isa ElementLevel
gpu_resident(A::Tensor) = separate_level(A) ? all(e -> on_gpu(e.val), A.lvl.lvl.val) :
                          on_gpu(values_of(A))
gpu_resident(S::SwizzleArray) = gpu_resident(S.body)
gpu_resident(x) = false
values_of(A::Tensor) = (l = A.lvl;
                        while !(l isa ElementLevel || l isa AtomicElementLevel); l = l.lvl; end;
                        l.val)

# A Dense(SparseList(Element)) or Dense(SparseDict(Element)) tensor with
# GPUVector buffers; swizzle(A, 2, 1) is the same descriptor read row-wise.
function descriptor(A::Tensor)
    sparse_level(A) || unsupported("operand is not Dense(SparseList/SparseDict(Element))")
    dense, sl = A.lvl, A.lvl.lvl
    el = sl.lvl
    if sl isa SparseDictLevel && !get(IN_ORDER, sl.val, false)
        unsupported("a SparseDict operand whose values are not in coordinate order")
    end
    all(on_gpu, (sl.ptr, sl.idx, el.val)) || unsupported("transfer(GPU(), A) the operand first")
    FgxCompressed(dense.shape, sl.shape, length(sl.idx), fgx_dtype(eltype(sl.ptr)),
                  fgx_dtype(eltype(sl.idx)), fgx_dtype(eltype(el.val)), 1,
                  pointer(sl.ptr), pointer(sl.idx), pointer(el.val))
end

"""Check a compressed operand against the SparseList invariants on the device
(fgx_validate): `FinchProtocolError` for a malformed `ptr` or an unsorted
fibre, `DimensionMismatch` for coordinates outside the shape."""
validate(dev::GPU, A) = check(ccall((:fgx_validate, libfgx), Cint,
                                    (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}),
                                    dev.ctx, descriptor(A), C_NULL))

device_of(A::Tensor) = values_of(A).dev
device_of(S::SwizzleArray) = device_of(S.body)

# ---- entry points ---------------------------------------------------------------------
# One wrapper per C entry point, over explicit device buffers.

# y .= 0; for j = _, i = _; y[i] += A[i, j] * x[j]   (test/reference64/typical/typical_spmv_csc.txt)
function spmv_csc!(dev::GPU, y::GPUVector, A::Tensor, x::GPUVector)
    check(ccall((:fgx_spmv_csc, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, descriptor(A), x.ptr, y.ptr, C_NULL))
    y
end

# The same y, run-to-run deterministic: a device transpose, then row sums.
function spmv_csc_ordered!(dev::GPU, y::GPUVector, A::Tensor, x::GPUVector)
    check(ccall((:fgx_spmv_csc_ordered, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, descriptor(A), x.ptr, y.ptr, C_NULL))
    y
end

# y .= 0; for i = _, j = _; y[i] += A[j, i] * x[j]: A's fibres are rows
# (benchmark/benchmarks.jl:365-400).
function spmv_csr!(dev::GPU, y::GPUVector, A::Tensor, x::GPUVector)
    check(ccall((:fgx_spmv_csr, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, descriptor(A), x.ptr, y.ptr, C_NULL))
    y
end

spmv!(dev::GPU, y::GPUVector, A::Tensor, x::GPUVector) = spmv_csc!(dev, y, A, x)
spmv!(dev::GPU, y::GPUVector, S::SwizzleArray{(2, 1)}, x::GPUVector) = spmv_csr!(dev, y, S.body, x)

# The same CSR SpMV through the entry point that allocates nothing, for code
# under HIP graph capture or loops that may not allocate; its scratch is a
# fixed per-GPU array of FGX_NOALLOC_SCRATCH_BYTES, shared by every call.
function spmv_noalloc!(dev::GPU, y::GPUVector, S::SwizzleArray{(2, 1)}, x::GPUVector;
                       stream::Ptr{Cvoid} = C_NULL)
    check(ccall((:fgx_spmv_csr_noalloc, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, descriptor(S.body), x.ptr, y.ptr, stream))
    y
end

# C[i, k] += A[i, j] * B[j, k] with A's fibres rows; B (n x k) and C (m x k)
# column-major, Finch's Dense(Dense(Element)) layout, leading dimensions n, m.
function spmm_csr!(dev::GPU, C::GPUVector, A::Tensor, B::GPUVector, k::Integer)
    d = descriptor(A)
    check(ccall((:fgx_spmm_csr, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Int64, Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}, Int64,
                 Cint, Ptr{Cvoid}),
                dev.ctx, d, k, B.ptr, d.n_inner, FGX_COL_MAJOR, C.ptr, d.n_outer, FGX_COL_MAJOR,
                C_NULL))
    C
end
spmm!(dev::GPU, C::GPUVector, S::SwizzleArray{(2, 1)}, B::GPUVector, k::Integer) =
    spmm_csr!(dev, C, S.body, B, k)

# out[q] = A[q] * dot(X[r, :], Y[:, c]) on A's pattern, A's fibres rows r:
# X is (rows x k) column-major; Y is (k x cols) column-major, i.e. the
# (cols x k) row-major factor the kernel reads.
function sddmm_csr!(dev::GPU, out::GPUVector, A::Tensor, X::GPUVector, ldx::Integer, xl::Int32,
                    Y::GPUVector, ldy::Integer, yl::Int32, k::Integer)
    check(ccall((:fgx_sddmm_csr, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Int64, Ptr{Cvoid}, Int64, Cint, Ptr{Cvoid}, Int64,
                 Cint, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, descriptor(A), k, X.ptr, ldx, xl, Y.ptr, ldy, yl, out.ptr, C_NULL))
    out
end

# Two-phase sparse outputs: symbolic -> size the buffers -> numeric, as the
# SparseList level's assemble/freeze protocol (sparse_list_levels.jl:301-327).
const EWISE_OPS = Dict("mul" => 0, "add" => 1, "max" => 2)
const SEMIRINGS = Dict("plus_times" => 0, "min_plus" => 1)

function ewise_buffers(dev::GPU, op::AbstractString, A::Tensor, B::Tensor)
    opc = EWISE_OPS[op]
    da, db = descriptor(A), descriptor(B)
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(dev, da.n_outer + 1)
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_ewise_symbolic, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ref{Int64},
                 Ptr{Cvoid}),
                dev.ctx, opc, da, db, ptr.ptr, nnz, C_NULL))
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(dev, nnz[])
    val = GPUVector{eltype(values_of(A))}(dev, nnz[])
    check(ccall((:fgx_ewise_numeric, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, opc, da, db, ptr.ptr, idx.ptr, val.ptr, C_NULL))
    (ptr, idx, val, da.n_inner, da.n_outer)
end

"""Both phases in one stream-ordered call (no host synchronisation), into
idx/val buffers of the bound's length (∩: min(nnz(A), nnz(B)); ∪: their sum);
nnz(C) = ptr[end] - 1 once the stream reaches it.  For stream-ordered
pipelines and graph capture."""
function ewise_async_buffers(dev::GPU, op::AbstractString, A::Tensor, B::Tensor)
    opc = EWISE_OPS[op]
    da, db = descriptor(A), descriptor(B)
    cap = op == "mul" ? min(da.nnz, db.nnz) : da.nnz + db.nnz
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(dev, da.n_outer + 1)
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(dev, cap)
    val = GPUVector{eltype(values_of(A))}(dev, cap)
    check(ccall((:fgx_ewise_async, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                dev.ctx, opc, da, db, ptr.ptr, idx.ptr, val.ptr, cap, C_NULL))
    (ptr, idx, val, da.n_inner, da.n_outer)
end

# Gustavson over fibre-major operands: fibre r of the result is
# sum over stored k of fibre r of A of A[k, r] * (fibre k of B)
# (docs/examples/spgemm.jl:59-76 in its row form).
function spgemm_buffers(dev::GPU, semiring::AbstractString, A::Tensor, B::Tensor)
    sr = SEMIRINGS[semiring]
    da, db = descriptor(A), descriptor(B)
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(dev, da.n_outer + 1)
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_spgemm_symbolic, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ref{Int64},
                 Ptr{Cvoid}),
                dev.ctx, sr, da, db, ptr.ptr, nnz, C_NULL))
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(dev, nnz[])
    val = GPUVector{eltype(values_of(A))}(dev, nnz[])
    check(ccall((:fgx_spgemm_numeric, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ref{FgxCompressed}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, sr, da, db, ptr.ptr, idx.ptr, val.ptr, C_NULL))
    (ptr, idx, val, db.n_inner, da.n_outer)
end

# The same matrix stored along its other dimension: permutedims of a
# Dense(SparseList) tensor (src/interface/copy.jl:64-71), and the transpose
# concordize inserts (src/scheduler/optimize.jl:396-452).
function transpose_buffers(dev::GPU, A::Tensor)
    d = descriptor(A)
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(dev, d.n_inner + 1)
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(dev, d.nnz)
    val = GPUVector{eltype(values_of(A))}(dev, d.nnz)
    check(ccall((:fgx_transpose, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                dev.ctx, d, ptr.ptr, idx.ptr, val.ptr, C_NULL))
    (ptr, idx, val, d.n_outer, d.n_inner)
end

# Positions 1:n of a frozen SparseDict whose entries were inserted fibre by
# fibre in coordinate order (sparse_dict_levels.jl:356-400), on the device.
function positions(dev::GPU, ::Type{Ti}, n::Integer) where {Ti}
    pos = GPUVector{Ti}(dev, n)
    check(ccall((:fgx_sequence, libfgx), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}),
                dev.ctx, fgx_dtype(Ti), pos.ptr, n, 1, C_NULL))
    IN_ORDER[pos] = true
    pos
end

# ---- result tensors ---------------------------------------------------------------
fill_of(semiring) = semiring == "min_plus" ? Inf : 0

function list_tensor((ptr, idx, val, n_inner, n_outer), fill)
    Tv = eltype(val)
    Tensor(DenseLevel(SparseListLevel{eltype(idx)}(ElementLevel{convert(Tv, fill),Tv}(val),
                                                   n_inner, ptr, idx), n_outer))
end

function dict_tensor(dev::GPU, (ptr, idx, val, n_inner, n_outer), fill)
    Tv, Ti = eltype(val), eltype(idx)
    pos = positions(dev, Int, length(idx))
    lvl = SparseDictLevel{Ti}(ElementLevel{convert(Tv, fill),Tv,Int}(val), n_inner, ptr, idx, pos,
                              Dict{Tuple{Int,Ti},Int}(), Int[])
    Tensor(DenseLevel(lvl, n_outer))
end

dense_vector(y::GPUVector{T}) where {T} = Tensor(DenseLevel(ElementLevel{zero(T),T}(y), length(y)))
dense_matrix(C::GPUVector{T}, m, k) where {T} =
    Tensor(DenseLevel(DenseLevel(ElementLevel{zero(T),T}(C), m), k))

"""permutedims(A) of a GPU-resident Dense(SparseList) matrix, built on the device."""
function Base.permutedims(dev::GPU, A::Tensor)
    list_tensor(transpose_buffers(dev, A), Finch.fill_value(A))
end

# Tensor(Dense(SparseList(Element(0))), fsparse(I, J, V, (m, n))) assembled on
# the device (src/interface/fsparse.jl:64-97): column fibres, rows sorted,
# duplicates summed left to right in input order.
function fsparse_csc(dev::GPU, I::GPUVector{Ti}, J::GPUVector{Ti}, V::GPUVector{Tv},
                     shape::Tuple{Integer,Integer}) where {Ti,Tv}
    m, n = shape
    ptr = GPUVector{Ti}(dev, n + 1)
    idx = GPUVector{Ti}(dev, length(V))
    val = GPUVector{Tv}(dev, length(V))
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_coo_to_compressed, libfgx), Cint,
                (Ptr{Cvoid}, Int64, Int64, Int64, Cint, Cint, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Int64}, Ptr{Cvoid}),
                dev.ctx, n, m, length(V), fgx_dtype(Ti), fgx_dtype(Tv), 1, J.ptr, I.ptr, V.ptr,
                ptr.ptr, idx.ptr, val.ptr, nnz, C_NULL))
    resize!(idx, nnz[]); resize!(val, nnz[])       # freeze: shrink to the distinct count
    list_tensor((ptr, idx, val, m, n), zero(Tv))
end

# dropfills(A) (src/interface/copy.jl:73-117) into the same format.
function Finch.dropfills(dev::GPU, A::Tensor)
    d = descriptor(A)
    ptr = GPUVector{eltype(A.lvl.lvl.ptr)}(dev, d.n_outer + 1)
    idx = GPUVector{eltype(A.lvl.lvl.idx)}(dev, d.nnz)
    val = GPUVector{eltype(values_of(A))}(dev, d.nnz)
    nnz = Ref{Int64}(0)
    check(ccall((:fgx_dropfills, libfgx), Cint,
                (Ptr{Cvoid}, Ref{FgxCompressed}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                 Ref{Int64}, Ptr{Cvoid}),
                dev.ctx, d, Float64(Finch.fill_value(A)), ptr.ptr, idx.ptr, val.ptr, nnz, C_NULL))
    resize!(idx, nnz[]); resize!(val, nnz[])
    list_tensor((ptr, idx, val, d.n_inner, d.n_outer), Finch.fill_value(A))
end

# The BF16/F16 storage form of an operand's values (FP32 accumulation in every
# kernel).  Finch has no 16-bit float element in this path; the converted
# buffer backs an ElementLevel whose values the kernels widen on load.
function convert_values!(dev::GPU, dst::GPUVector, dst_type::Int32, src::GPUVector, src_type::Int32)
    check(ccall((:fgx_convert, libfgx), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cint, Int64, Ptr{Cvoid}),
                dev.ctx, src.ptr, src_type, dst.ptr, dst_type, length(src), C_NULL))
    dst
end

# ---- the routing table ----------------------------------------------------------------
const ROUTES = TOML.parsefile(ROUTES_FILE)["route"]

"""An operand of a program: the tensor bound to it and the indices it is read
with, plus its kind and whether its storage order is the reverse of the read."""
struct Operand
    tns::Any
    idxs::Vector{Symbol}
end

kind_of(A::Tensor) = sparse_level(A) ? ("sparse", false) :
                     dense_level(A) ? ("dense", false) : unsupported("operand format $(summary(A))")
kind_of(S::SwizzleArray{(2, 1)}) = sparse_level(S.body) ? ("sparse", true) :
                                   unsupported("a swizzled operand that is not sparse")
kind_of(x) = unsupported("operand $(typeof(x))")
storage_tensor(A::Tensor) = A
storage_tensor(S::SwizzleArray) = S.body

"""
    canonical(out_idxs, reduce, operands, combine) -> (program, kinds)

The canonical form routes.toml lists: sparse operands written in storage
order (a swizzled or transposed read reversed), tensors renamed C, A, B, D,
indices renamed i, j, k, l by first appearance (output first)."""
function canonical(out_idxs::Vector{Symbol}, reduce, operands::Vector{Operand}, combine::String)
    names = Dict{Symbol,Char}()
    nm(x) = get!(names, x, "ijklmn"[length(names) + 1])
    out = [nm(x) for x in out_idxs]
    kinds = String[]
    accs = String[]
    for (n, o) in enumerate(operands)
        kind, rev = kind_of(o.tns)
        push!(kinds, kind)
        idxs = kind == "sparse" && rev ? reverse(o.idxs) : o.idxs
        push!(accs, string("ABDE"[n], "[", join([nm(x) for x in idxs], ","), "]"))
    end
    assign = reduce === nothing ? "=" : reduce == "+" ? "+=" : "<<$(reduce)>>="
    rhs = combine == "max" ? "max($(accs[1]), $(accs[2]))" : join(accs, " $combine ")
    ("C[$(join(out, ","))] $assign $rhs", kinds)
end

function lookup(program::String, kinds::Vector{String})
    for r in ROUTES
        r["program"] == program && r["operands"] == kinds && return r
    end
    nothing
end

# ---- handlers: one per routes.toml `handler` ------------------------------------------
# Each takes the route and the operand tensors in program order and returns
# (result, storage): storage "csc"/"csr" for sparse buffers (fibres along
# columns / rows), "dense" for a dense Tensor.

col_vector(x::Tensor) = values_of(x)
dims_of(B::Tensor) = (B.lvl.lvl.shape, B.lvl.shape)      # Dense(Dense): rows, columns

function route_spmv_csc(r, (A, x), dev)
    At = storage_tensor(A)
    y = GPUVector{eltype(values_of(x))}(dev, At.lvl.lvl.shape)
    (dense_vector(spmv_csc!(dev, y, At, col_vector(x))), "dense")
end
function route_spmv_csr(r, (A, x), dev)
    At = storage_tensor(A)
    y = GPUVector{eltype(values_of(x))}(dev, At.lvl.shape)
    (dense_vector(spmv_csr!(dev, y, At, col_vector(x))), "dense")
end
function route_spmm_csr(r, (A, B), dev; At = storage_tensor(A))
    n, k = dims_of(B)
    m = At.lvl.shape
    C = GPUVector{eltype(values_of(B))}(dev, m * k)
    (dense_matrix(spmm_csr!(dev, C, At, values_of(B), k), m, k), "dense")
end
route_spmm_csc(r, (A, B), dev) =
    route_spmm_csr(r, (A, B), dev; At = list_tensor(transpose_buffers(dev, storage_tensor(A)), 0))
function sddmm_on(dev, A::Tensor, F1::Tensor, F2::Tensor, fibres_are_rows::Bool)
    (m1, k) = dims_of(F1)                      # X: rows x k
    (k2, n2) = dims_of(F2)                     # Y: k x cols
    out = GPUVector{eltype(values_of(A))}(dev, length(A.lvl.lvl.idx))
    if fibres_are_rows      # X rows by A's fibres, Y columns by A's coordinates
        sddmm_csr!(dev, out, A, values_of(F1), m1, FGX_COL_MAJOR, values_of(F2), k2,
                   FGX_ROW_MAJOR, k)
    else                    # the factors' roles exchanged on column fibres
        sddmm_csr!(dev, out, A, values_of(F2), k2, FGX_ROW_MAJOR, values_of(F1), m1,
                   FGX_COL_MAJOR, k)
    end
    sl = A.lvl.lvl
    ((sl.ptr, sl.idx, out, sl.shape, A.lvl.shape), fibres_are_rows ? "csr" : "csc")
end
route_sddmm_csr(r, (M, X, Y), dev) = sddmm_on(dev, storage_tensor(M), X, Y, true)
route_sddmm_csc(r, (M, X, Y), dev) = sddmm_on(dev, storage_tensor(M), X, Y, false)
route_ewise(r, (A, B), dev) =
    (ewise_buffers(dev, r["op"], storage_tensor(A), storage_tensor(B)), "csc")
route_ewise_rows(r, (A, B), dev) =
    (ewise_buffers(dev, r["op"], storage_tensor(A), storage_tensor(B)), "csr")
other_order(dev, B) = list_tensor(transpose_buffers(dev, storage_tensor(B)), 0)
route_ewise_transpose_b(r, (A, B), dev) =
    (ewise_buffers(dev, r["op"], storage_tensor(A), other_order(dev, B)), "csc")
route_ewise_rows_transpose_b(r, (A, B), dev) =
    (ewise_buffers(dev, r["op"], storage_tensor(A), other_order(dev, B)), "csr")
# Column-compressed C = A*B column by column is (B^T A^T)^T.
route_spgemm_csc(r, (A, B), dev) =
    (spgemm_buffers(dev, r["semiring"], storage_tensor(B), storage_tensor(A)), "csc")
route_spgemm_csr(r, (A, B), dev) =
    (spgemm_buffers(dev, r["semiring"], storage_tensor(A), storage_tensor(B)), "csr")
route_spgemm_csr_transpose_b(r, (A, B), dev) =
    (spgemm_buffers(dev, r["semiring"], storage_tensor(A), other_order(dev, B)), "csr")
route_spgemm_csc_transpose_b(r, (A, B), dev) =
    (spgemm_buffers(dev, r["semiring"], other_order(dev, B), storage_tensor(A)), "csc")

const HANDLERS = Dict{String,Function}(
    "spmv_csc" => route_spmv_csc, "spmv_csr" => route_spmv_csr,
    "spmm_csr" => route_spmm_csr, "spmm_csc" => route_spmm_csc,
    "sddmm_csr" => route_sddmm_csr, "sddmm_csc" => route_sddmm_csc,
    "ewise" => route_ewise, "ewise_rows" => route_ewise_rows,
    "ewise_transpose_b" => route_ewise_transpose_b,
    "ewise_rows_transpose_b" => route_ewise_rows_transpose_b,
    "spgemm_csc" => route_spgemm_csc, "spgemm_csr" => route_spgemm_csr,
    "spgemm_csr_transpose_b" => route_spgemm_csr_transpose_b,
    "spgemm_csc_transpose_b" => route_spgemm_csc_transpose_b,
)

"""Run a routed program; the result as (value, storage)."""
function run_route(out_idxs, reduce, operands::Vector{Operand}, combine; debug::Bool = false)
    program, kinds = canonical(out_idxs, reduce, operands, combine)
    r = lookup(program, kinds)
    r === nothing && unsupported("program $program over $kinds has no device route")
    dev = device_of(first(o.tns for o in operands if gpu_resident(o.tns)))
    if debug    # Finch's mode=:debug: check every sparse operand first
        for o in operands
            kind_of(o.tns)[1] == "sparse" && validate(dev, storage_tensor(o.tns))
        end
    end
    res, storage = HANDLERS[r["handler"]](r, Tuple(o.tns for o in operands), dev)
    (r, dev, res, storage)
end

# ---- @finch: the hook on Finch.execute ---------------------------------------------------
# `@finch` builds a program instance and calls `Finch.execute(prgm; mode)`
# (src/execute.jl:180-220).  Programs whose tensors are GPU-resident are
# routed here; every other program keeps Finch's staged `execute_impl`.

uses_gpu(ex::TagInstance) = gpu_resident(ex.bind)
uses_gpu(ex::FinchNodeInstance) = any(uses_gpu, instance_children(ex))
uses_gpu(ex) = false
instance_children(ex::BlockInstance) = ex.bodies
instance_children(ex::LoopInstance) = (ex.body,)
instance_children(ex::AssignInstance) = (ex.lhs, ex.rhs)
instance_children(ex::AccessInstance) = (ex.tns,)
instance_children(ex::CallInstance) = ex.args
instance_children(ex::DeclareInstance) = (ex.tns,)
instance_children(ex::FreezeInstance) = (ex.tns,)
instance_children(ex::ThawInstance) = (ex.tns,)
instance_children(ex::FinchNodeInstance) = ()

# The single assignment of a program, the loops around it, and its declares.
function statements!(acc, ex::BlockInstance)
    foreach(b -> statements!(acc, b), ex.bodies)
    acc
end
statements!(acc, ex::LoopInstance) = statements!(acc, ex.body)
statements!(acc, ex::AssignInstance) = (push!(acc, ex); acc)
statements!(acc, ex::Union{DeclareInstance,FreezeInstance,ThawInstance,YieldBindInstance}) = acc
statements!(acc, ex) = unsupported("program node $(typeof(ex)) is not routed")

idx_name(i::IndexInstance) = i.name
idx_name(i) = unsupported("index expression $(typeof(i))")
operand(a::AccessInstance) = Operand(a.tns.bind, [idx_name(i) for i in a.idxs])
operand(a) = unsupported("right-hand term $(typeof(a))")
literal(x::LiteralInstance) = x.val
literal(x) = unsupported("operator $(typeof(x))")

function reduce_name(op)
    op === (+) && return "+"
    op === min && return "min"
    op === max && return "max"
    op isa Finch.FinchNotation.InitWriter && return nothing
    op === Finch.FinchNotation.overwrite && return nothing
    unsupported("reduction $op")
end
combine_name(op) = op === (*) ? "*" : op === (+) ? "+" : op === max ? "max" :
                   unsupported("operator $op")

yield_names(ex::BlockInstance) = vcat([yield_names(b) for b in ex.bodies]...)
yield_names(ex::YieldBindInstance) = [a.name for a in ex.args]
yield_names(ex) = Symbol[]

function Finch.execute(ex::BlockInstance; algebra = Finch.DefaultAlgebra(), mode = :safe)
    uses_gpu(ex) || return Finch.execute_impl(ex, Val(algebra), Val(mode))
    algebra == Finch.DefaultAlgebra() || unsupported("a custom algebra on GPU operands")
    stmts = statements!(AssignInstance[], ex)
    length(stmts) == 1 || unsupported("programs with $(length(stmts)) assignments")
    st = stmts[1]
    lhs = st.lhs
    out_tns = lhs.tns.bind
    out_idxs = [idx_name(i) for i in lhs.idxs]
    reduce = reduce_name(literal(st.op))
    rhs = st.rhs
    rhs isa CallInstance || unsupported("right-hand side $(typeof(rhs))")
    combine = combine_name(literal(rhs.op))
    operands = Operand[operand(a) for a in rhs.args]
    r, dev, res, storage = run_route(out_idxs, reduce, operands, combine;
                                     debug = mode === :debug)
    write_output!(out_tns, r, dev, res, storage)
    names = yield_names(ex)
    isempty(names) && (names = [lhs.tns.var.name])
    NamedTuple{Tuple(names)}(Tuple(n == lhs.tns.var.name ? out_tns :
                                   unsupported("yielding an input") for n in names))
end

# The declared output keeps its format: a dense output is overwritten with
# the result, a Dense(SparseList) output takes the result's buffers (its
# resize! and freeze), transposed first when its storage order differs.
function write_output!(out::Tensor, r, dev, res, storage)
    if storage == "dense" && separate_level(out)
        # One copy into each position's own level (Separate gives every
        # position a separate allocation, so there is no single buffer).
        src, lvls = values_of(res), out.lvl.lvl.val
        length(src) == length(lvls) || throw(DimensionMismatch("output has $(length(lvls)) " *
                                                               "positions, result $(length(src))"))
        T = eltype(src)
        for (p, e) in enumerate(lvls)
            on_gpu(e.val) && length(e.val) == 1 && eltype(e.val) == T ||
                unsupported("position $p of the Separate output is not a GPU value of type $T")
            hip_copy(e.val.ptr, src.ptr + (p - 1) * sizeof(T), sizeof(T), D2D)
        end
        return out
    end
    if storage == "dense"
        dense_level(out) || unsupported("the output of $(r["program"]) must be dense")
        src, dst = values_of(res), values_of(out)
        length(src) == length(dst) || throw(DimensionMismatch("output has $(length(dst)) " *
                                                              "entries, result $(length(src))"))
        hip_copy(dst.ptr, src.ptr, length(src) * sizeof(eltype(src)), D2D)
        return out
    end
    (out.lvl isa DenseLevel && out.lvl.lvl isa SparseListLevel) ||
        unsupported("the output of $(r["program"]) must be Dense(SparseList)")
    storage == "csc" || (res = transpose_buffers(dev, list_tensor(res, 0)))
    ptr, idx, val, n_inner, n_outer = res
    sl, el = out.lvl.lvl, out.lvl.lvl.lvl
    (sl.shape, out.lvl.shape) == (n_inner, n_outer) ||
        throw(DimensionMismatch("output is $(sl.shape)×$(out.lvl.shape), result $(n_inner)×$(n_outer)"))
    for (dst, src) in ((sl.ptr, ptr), (sl.idx, idx), (el.val, val))
        resize!(dst, length(src))
        hip_copy(dst.ptr, src.ptr, length(src) * sizeof(eltype(src)), D2D)
    end
    out
end
function write_output!(out::SwizzleArray{(2, 1)}, r, dev, res, storage)
    storage == "dense" && unsupported("a swizzled dense output")
    # The body stores the result's rows as its fibres.
    rows = storage == "csr" ? res : transpose_buffers(dev, list_tensor(res, 0))
    write_output!(out.body, r, dev, rows, "csc")
    out
end
write_output!(out, r, dev, res, storage) = unsupported("output $(typeof(out))")

# ---- compute: the scheduler ------------------------------------------------------------
# compute(lazy(...)) hands the scheduler a logic program
# plan(query(alias, expr)..., produces(aliases...)) (src/interface/lazy.jl:715-730).
# GPUScheduler reads each query's expression as an einsum over leaves and runs
# its route; programs over host tensors go to the fallback scheduler.

"""
    GPUScheduler(; fallback = Finch.default_scheduler())

A `compute` scheduler for GPU-resident operands; install it with
`Finch.set_scheduler!(FinchGPU.GPUScheduler())` or scope it with
`Finch.with_scheduler(f, FinchGPU.GPUScheduler())`.  Sparse results are
`Dense(SparseDict)` tensors, as Finch's scheduler returns them.
"""
struct GPUScheduler
    fallback::Any
    verbose::Bool
end
GPUScheduler(; fallback = Finch.default_scheduler(), verbose = false) =
    GPUScheduler(fallback, verbose)
Finch.set_options(s::GPUScheduler; verbose = s.verbose, kwargs...) =
    GPUScheduler(Finch.set_options(s.fallback; verbose = verbose, kwargs...), verbose)

# An expression read as an einsum term tree: leaves with field names, maps,
# and reductions.
abstract type LTerm end
struct LLeaf <: LTerm
    tns::Any
    fields::Vector{Symbol}
end
struct LMap <: LTerm
    op::Any
    args::Vector{LTerm}
end
struct LAgg <: LTerm
    op::Any
    arg::LTerm
    fields::Vector{Symbol}
end
# The same term with its output fields in another order (no data moves).
struct LReorder <: LTerm
    arg::LTerm
    fields::Vector{Symbol}
end

# Output fields in order: a relabel renames them position by position.
out_fields(t::LLeaf) = t.fields
out_fields(t::LMap) = unique(vcat([out_fields(a) for a in t.args]...))
out_fields(t::LAgg) = setdiff(out_fields(t.arg), t.fields)
out_fields(t::LReorder) = t.fields

rename(t::LLeaf, m) = LLeaf(t.tns, [get(m, f, f) for f in t.fields])
rename(t::LMap, m) = LMap(t.op, LTerm[rename(a, m) for a in t.args])
rename(t::LAgg, m) = LAgg(t.op, rename(t.arg, m), [get(m, f, f) for f in t.fields])
rename(t::LReorder, m) = LReorder(rename(t.arg, m), [get(m, f, f) for f in t.fields])

unwrap(t::LReorder) = unwrap(t.arg)
unwrap(t::LTerm) = t

fieldnames_of(nodes) = Symbol[n.name for n in nodes]
imm(n::LogicNode) = n.kind === immediate ? n.val : unsupported("logic operator $(n.kind)")

function lterm(n::LogicNode, env)
    k = n.kind
    if k === subquery
        t = lterm(n.arg, env)
        env[n.lhs.name] = t
        return t
    elseif k === alias
        haskey(env, n.name) || unsupported("alias $(n.name) defined outside the plan")
        return env[n.name]
    elseif k === table
        return LLeaf(imm(n.tns), fieldnames_of(n.idxs))
    elseif k === relabel
        t = lterm(n.arg, env)
        old = out_fields(t)
        new = fieldnames_of(n.idxs)
        length(old) == length(new) || unsupported("relabel of $(length(old)) fields")
        return rename(t, Dict(zip(old, new)))
    elseif k === reorder
        t = lterm(n.arg, env)
        f = fieldnames_of(n.idxs)
        Set(out_fields(t)) == Set(f) || unsupported("broadcast by reorder")
        return LReorder(t, f)
    elseif k === mapjoin
        return LMap(imm(n.op), LTerm[lterm(a, env) for a in n.args])
    elseif k === aggregate
        return LAgg(imm(n.op), lterm(n.arg, env), fieldnames_of(n.idxs))
    elseif k === reformat
        return lterm(n.arg, env)
    end
    unsupported("logic node $k")
end

leaf_operand(t::LReorder) = leaf_operand(t.arg)
leaf_operand(t::LLeaf) = Operand(t.tns, t.fields)
leaf_operand(t) = unsupported("a nested expression where a tensor is expected")

"""The einsum a query's expression computes: (output fields, reduce, operands,
combine).  Each leaf keeps the fields of its own storage order."""
function as_einsum(t0::LTerm)
    out = out_fields(t0)
    t = unwrap(t0)
    if t isa LAgg && unwrap(t.arg) isa LMap                   # C[..] <<op>>= A ⊗ B
        m = unwrap(t.arg)
        return (out, reduce_name(t.op), Operand[leaf_operand(a) for a in m.args],
                combine_name(m.op))
    elseif t isa LMap && length(t.args) == 2 && unwrap(t.args[2]) isa LAgg &&
           unwrap(unwrap(t.args[2]).arg) isa LMap && t.op === (*)   # M .* (X * Y): SDDMM
        a = unwrap(unwrap(t.args[2]).arg)
        return (out, "+", Operand[leaf_operand(t.args[1]);
                                  [leaf_operand(x) for x in a.args]], "*")
    elseif t isa LMap                                         # C[..] = A op B
        return (out, nothing, Operand[leaf_operand(a) for a in t.args], combine_name(t.op))
    end
    unsupported("expression $(typeof(t))")
end

logic_uses_gpu(n::LogicNode) = n.kind === immediate ? gpu_resident(n.val) :
                               any(logic_uses_gpu, n.children)

function (s::GPUScheduler)(prgm::LogicNode)
    logic_uses_gpu(prgm) || return s.fallback(prgm)
    prgm.kind === plan || unsupported("logic program $(prgm.kind)")
    env = Dict{Symbol,Any}()
    results = Dict{Symbol,Any}()
    outs = Symbol[]
    for body in prgm.bodies
        if body.kind === query
            t = lterm(body.rhs, env)
            out, reduce, operands, combine = as_einsum(t)
            r, dev, res, storage = run_route(out, reduce, operands, combine)
            if storage == "dense"
                results[body.lhs.name] = res
                env[body.lhs.name] = LLeaf(res, out)
            else
                c = dict_tensor(dev, res, fill_of(get(r, "semiring", nothing)))
                # A result stored by rows is C's transpose in storage order.
                results[body.lhs.name] = storage == "csr" ? Finch.swizzle(c, 2, 1) : c
                env[body.lhs.name] = LLeaf(c, storage == "csr" ? reverse(out) : out)
            end
        elseif body.kind === produces
            outs = Symbol[a.name for a in body.args]
        else
            unsupported("plan statement $(body.kind)")
        end
    end
    Tuple(results[o] for o in outs)
end

# ---- batched form ---------------------------------------------------------------------
struct FgxBatchItem           # struct fgx_batch_item
    A::FgxCompressed
    n_rhs::Int64
    B::Ptr{Cvoid}
    C::Ptr{Cvoid}
end

const FGX_BATCH_FORCE_ROWSPLIT, FGX_BATCH_FORCE_NNZSPLIT, FGX_BATCH_NO_GRAPH,
      FGX_BATCH_BF16_INPUTS, FGX_BATCH_MEASURE = 1, 2, 4, 8, 16

"""Many independent `C = A*B` (CSR A, row-major F32 B/C with `n_rhs` columns)
in one persistent launch, replayed from a hipGraph.  The items' buffers must
outlive the plan.  With `flags = FGX_BATCH_BF16_INPUTS` every B is BF16
(the items' values BF16 too); C stays F32."""
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
    input_type::Int32         # dense inputs: 0 / FGX_F32, or FGX_BF16 (batched, BF16 values)
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
               max_batch::Integer = 512, plain::Bool = false, bf16_inputs::Bool = false)
    cfg = FgxServeConfig(hbm_budget_bytes, max_slots, max_batch, plain ? 1 : 0,
                         bf16_inputs ? FGX_BF16 : Int32(0))
    res = Vector{FgxRequestResult}(undef, length(reqs))
    check(ccall((:fgx_serve_run, libfgx), Cint,
                (Ptr{Cvoid}, Ptr{FgxCompressed}, Int64, Ptr{FgxRequest}, Int64,
                 Ref{FgxServeConfig}, Ptr{FgxRequestResult}),
                dev.ctx, operands, length(operands), reqs, length(reqs), cfg, res))
    res
end

end # module
