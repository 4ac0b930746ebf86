# DaggerFFTHIP.jl -- Julia binding of libdaggerfft_hip (MI355X / gfx950).
#
# Drop-in for DaggerFFT's single-GPU transform path: the README API
# (fft/ifft/rfft/irfft with decomp=Pencil()/Slab(), dims, transforms=R2R(...),
# README.md:29-67) and the exported bang API (fft!/ifft!, src/DaggerFFT.jl:32-37)
# over ROCArrays, each call a ccall into the C ABI of include/daggerfft_hip.h.
# Arrays are Julia column-major, which is the library's layout; dims are
# 1-based here and 0-based in the C ABI.
module DaggerFFTHIP

using AMDGPU

export fft, ifft, rfft, irfft, fft!, ifft!, Pencil, Slab, R2R
export FFT, IFFT, RFFT, IRFFT, FFT!, IFFT!, RFFT!, IRFFT!
export create_workspace, cleanup_workspace!, plan_fft, plan_ifft, plan_rfft, plan_irfft
export BlockedArray, distribute, compute_overlap, redistribute!
export Batch, BatchDesc, Server, submit, wait_ticket, result!, release

const libdfft = joinpath(@__DIR__, "..", "libdaggerfft_hip.so")

const DFFT_C2C, DFFT_R2C, DFFT_C2R, DFFT_R2R = 0, 1, 2, 3
const DFFT_F32, DFFT_F64 = 0, 1
const DFFT_PENCIL, DFFT_SLAB = 0, 1
const DFFT_FLAG_NORMALIZE_R2R = UInt32(0x2)

# tag types (src/fft.jl:13-24)
abstract type Decomposition end
struct Pencil <: Decomposition end
struct Slab <: Decomposition end
struct FFT end; struct IFFT end; struct RFFT end; struct IRFFT end
struct FFT! end; struct IFFT! end; struct RFFT! end; struct IRFFT! end
struct R2R{N}
    kinds::NTuple{N,Int32}   # FFTW fftw_r2r_kind values (FFTW.REDFT10 == 5, ...)
end
R2R(k::Tuple) = R2R{length(k)}(Int32.(k))

decomp_code(::Pencil) = DFFT_PENCIL
decomp_code(::Slab) = DFFT_SLAB
prec_code(::Type{<:Union{Float64,ComplexF64}}) = DFFT_F64
prec_code(::Type{<:Union{Float32,ComplexF32}}) = DFFT_F32

struct DFFTError <: Exception
    status::Cint
end
Base.showerror(io::IO, e::DFFTError) =
    print(io, "libdaggerfft_hip: ", unsafe_string(ccall((:dfft_strerror, libdfft), Cstring, (Cint,), e.status)))
check(st) = st == 0 ? nothing : throw(DFFTError(st))

mutable struct Plan
    handle::Ptr{Cvoid}
    insize::Dims
    outsize::Dims
end

function Plan(sz::Dims, kind, sign, prec, decomp::Decomposition, dims; kinds=nothing, flags=UInt32(0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    n = Int64[sz...]
    d = Cint[(x - 1) for x in dims]
    k = kinds === nothing ? C_NULL : Cint[kinds...]
    check(ccall((:dfft_plan_create, libdfft), Cint,
                (Ptr{Ptr{Cvoid}}, Cint, Ptr{Int64}, Cint, Cint, Ptr{Cint}, Cint, Cint, Ptr{Cint}, Cint, UInt32),
                h, length(sz), n, kind, sign, k, prec, decomp_code(decomp), d, length(d), flags))
    p = Plan(h[], sz, sz)
    finalizer(p) do q
        q.handle != C_NULL && ccall((:dfft_plan_destroy, libdfft), Cint, (Ptr{Cvoid},), q.handle)
        q.handle = C_NULL
    end
    return p
end

# the plan cache (GPU_PLAN_CACHE, src/fftgpu.jl:83,107-116), under a lock
const PLAN_CACHE = Dict{Any,Plan}()
const PLAN_LOCK = ReentrantLock()
cached(key, mk) = lock(() -> get!(mk, PLAN_CACHE, key), PLAN_LOCK)

stream_ptr() = AMDGPU.stream().stream

function execute!(p::Plan, out::ROCArray, inp::ROCArray)
    check(ccall((:dfft_execute, libdfft), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                p.handle, pointer(inp), pointer(out), C_NULL, stream_ptr()))
    return out
end
LinearAlgebra_mul!(out, p::Plan, inp) = execute!(p, out, inp)

# execute! without any device allocation (for code under graph capture): a C2R
# plan's workspace is the library's static per-device scratch buffer
function execute_static!(p::Plan, out::ROCArray, inp::ROCArray)
    check(ccall((:dfft_execute_static, libdfft), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                p.handle, pointer(inp), pointer(out), stream_ptr()))
    return out
end

_dims(A, dims) = dims === nothing ? Tuple(1:ndims(A)) : Tuple(dims)

function r2r_kinds(transforms, dims, decomp)
    ts = transforms isa Tuple ? transforms : (transforms,)
    any(t -> t isa R2R, ts) || return nothing
    if length(ts) == 1
        k = ts[1].kinds
        return length(k) == 1 ? ntuple(_ -> k[1], length(dims)) : k
    end
    # Slab: transforms[1] covers the 2D stage, transforms[2] is ignored (src/fft.jl:735,746)
    decomp isa Slab && length(ts) == 3 && return (ts[1].kinds[1], ts[1].kinds[end], ts[3].kinds[1])
    return ntuple(i -> ts[i].kinds[1], length(dims))
end

plan_fft(A::ROCArray{T}, dims=nothing; decomp=Pencil()) where {T<:Complex} =
    cached((size(A), T, :fft, _dims(A, dims), decomp),
           () -> Plan(size(A), DFFT_C2C, -1, prec_code(T), decomp, _dims(A, dims)))
plan_ifft(A::ROCArray{T}, dims=nothing; decomp=Pencil()) where {T<:Complex} =
    cached((size(A), T, :ifft, _dims(A, dims), decomp),
           () -> Plan(size(A), DFFT_C2C, +1, prec_code(T), decomp, _dims(A, dims)))
plan_rfft(A::ROCArray{T}, dims=nothing; decomp=Pencil()) where {T<:Real} =
    cached((size(A), T, :rfft, _dims(A, dims), decomp),
           () -> Plan(size(A), DFFT_R2C, -1, prec_code(T), decomp, _dims(A, dims)))
function plan_irfft(F::ROCArray{T}, d::Integer, dims=nothing; decomp=Pencil()) where {T<:Complex}
    dd = _dims(F, dims)
    sz = ntuple(i -> i == dd[1] ? d : size(F, i), ndims(F))
    cached((sz, T, :irfft, dd, decomp), () -> Plan(sz, DFFT_C2R, +1, prec_code(T), decomp, dd))
end

# ---- README API ------------------------------------------------------------
function fft(A::ROCArray{T}; decomp=Pencil(), dims=nothing, transforms=nothing) where {T}
    dd = _dims(A, dims)
    ks = transforms === nothing ? nothing : r2r_kinds(transforms, dd, decomp)
    if ks !== nothing
        p = cached((size(A), T, :r2r, dd, ks, decomp),
                   () -> Plan(size(A), DFFT_R2R, -1, prec_code(T), decomp, dd; kinds=ks))
        return execute!(p, similar(A), A)
    end
    Ac = T <: Complex ? A : complex.(A)
    return execute!(plan_fft(Ac, dd; decomp), similar(Ac), Ac)
end

function ifft(F::ROCArray{T}; decomp=Pencil(), dims=nothing, transforms=nothing) where {T}
    dd = _dims(F, dims)
    ks = transforms === nothing ? nothing : r2r_kinds(transforms, dd, decomp)
    if ks !== nothing   # normalised so that REDFT01 after REDFT10 reconstructs the input
        p = cached((size(F), T, :ir2r, dd, ks, decomp),
                   () -> Plan(size(F), DFFT_R2R, +1, prec_code(T), decomp, dd; kinds=ks, flags=DFFT_FLAG_NORMALIZE_R2R))
        return execute!(p, similar(F), F)
    end
    return execute!(plan_ifft(F, dd; decomp), similar(F), F)
end

function rfft(A::ROCArray{T}; decomp=Pencil(), dims=nothing) where {T<:Real}
    dd = _dims(A, dims)
    osz = ntuple(i -> i == dd[1] ? size(A, i) ÷ 2 + 1 : size(A, i), ndims(A))
    return execute!(plan_rfft(A, dd; decomp), ROCArray{Complex{T}}(undef, osz), A)
end

function irfft(F::ROCArray{Complex{T}}, d::Integer; decomp=Pencil(), dims=nothing) where {T<:Real}
    dd = _dims(F, dims)
    p = plan_irfft(F, d, dd; decomp)
    return execute!(p, ROCArray{T}(undef, p.insize), F)
end

# ---- bang API over single-device arrays (src/fft.jl:651-776) ---------------
# The transforms tuple selects C2C or, with R2R entries, the r2r kinds per dim
# (Slab: transforms[1] covers the 2D stage, transforms[2] is ignored,
# src/fft.jl:735,746).
function _bang_plan(A::ROCArray{T}, transforms, dims, decomp, sign) where {T}
    ks = r2r_kinds(transforms, dims, decomp)
    if ks !== nothing
        flags = sign > 0 ? DFFT_FLAG_NORMALIZE_R2R : UInt32(0)
        return cached((size(A), T, sign > 0 ? :ir2r : :r2r, Tuple(dims), ks, decomp),
                      () -> Plan(size(A), DFFT_R2R, sign, prec_code(T), decomp, dims; kinds=ks, flags=flags))
    end
    return sign < 0 ? plan_fft(A, dims; decomp) : plan_ifft(A, dims; decomp)
end
struct Workspace
    src_size::Dims
    dst_size::Dims
end
create_workspace(src::ROCArray, dst::ROCArray; kwargs...) = Workspace(size(src), size(dst))
cleanup_workspace!(::Workspace) = lock(() -> empty!(PLAN_CACHE), PLAN_LOCK)

stage_inverse(transforms) = first(transforms) isa Union{IFFT,IFFT!}

# Pencil fft!(C, A, B, ...): input A, result C; ifft!(C, A, B, ...): input A
# (the spectrum), result C -- as the reference's stage chain A -> B -> C.
function fft!(C::ROCArray, A::ROCArray, B::ROCArray, ws_AB::Workspace, ws_BC::Workspace, transforms, dims)
    return execute!(_bang_plan(A, transforms, Tuple(dims), Pencil(), -1), C, A)
end
function fft!(B::ROCArray, A::ROCArray, ws::Workspace, transforms, dims, ::Slab)
    return execute!(_bang_plan(A, transforms, Tuple(dims), Slab(), -1), B, A)
end
function ifft!(C::ROCArray, A::ROCArray, B::ROCArray, ws_AB::Workspace, ws_BC::Workspace, transforms, dims)
    return execute!(_bang_plan(A, transforms, Tuple(dims), Pencil(), +1), C, A)
end
function ifft!(A::ROCArray, B::ROCArray, ws::Workspace, transforms, dims, ::Slab)
    return execute!(_bang_plan(B, transforms, Tuple(dims), Slab(), +1), A, B)
end

# ---- bang API over Blocks-distributed arrays on one device ----------------
# A BlockedArray is a DArray with Blocks(b1,b2,b3) whose chunks all live on
# this GPU.  create_workspace keeps the local patterns of the reference's
# create_workspace (src/fft.jl:233-343, the branch at :301-305),
# compute_overlap is src/fft.jl:176-178, every pattern is moved by
# dfft_box_copy (local_copy!, src/fft.jl:216-231), and the stage drivers
# follow src/fft.jl:651-776 with per-chunk in-place transforms (apply_fft!).
struct BlockedArray{T}
    size::NTuple{3,Int}
    subdomains::Vector{NTuple{3,UnitRange{Int}}}
    chunks::Vector{ROCArray{T,3}}
end

function BlockedArray{T}(sz::NTuple{3,Int}, blocks::NTuple{3,Int}) where {T}
    b = ntuple(i -> clamp(blocks[i], 1, sz[i]), 3)
    doms = NTuple{3,UnitRange{Int}}[]
    for k0 in 1:b[3]:sz[3], j0 in 1:b[2]:sz[2], i0 in 1:b[1]:sz[1]   # dim 1 fastest
        push!(doms, (i0:min(i0 + b[1] - 1, sz[1]), j0:min(j0 + b[2] - 1, sz[2]), k0:min(k0 + b[3] - 1, sz[3])))
    end
    return BlockedArray{T}(sz, doms, [ROCArray{T}(undef, length.(d)) for d in doms])
end

# dst[dlo .+ (0:cnt-1)...] = src[slo .+ (0:cnt-1)...]; 1-based lo as in Julia
function box_copy!(dst::ROCArray{T,3}, dlo, src::ROCArray{T,3}, slo, cnt) where {T}
    any(==(0), cnt) && return dst
    check(ccall((:dfft_box_copy, libdfft), Cint,
                (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Cint, Ptr{Cvoid}),
                pointer(dst), Int64[size(dst)...], Int64[(dlo .- 1)...], pointer(src), Int64[size(src)...],
                Int64[(slo .- 1)...], Int64[cnt...], sizeof(T), stream_ptr()))
    return dst
end

function distribute(A::ROCArray{T,3}, blocks::NTuple{3,Int}) where {T}
    D = BlockedArray{T}(size(A), blocks)
    for (d, ch) in zip(D.subdomains, D.chunks)
        box_copy!(ch, (1, 1, 1), A, first.(d), length.(d))
    end
    return D
end

function Base.collect(D::BlockedArray{T}) where {T}
    out = ROCArray{T}(undef, D.size)
    for (d, ch) in zip(D.subdomains, D.chunks)
        box_copy!(out, first.(d), ch, (1, 1, 1), length.(d))
    end
    return out
end

compute_overlap(src_domain, dst_domain) = ntuple(i -> intersect(src_domain[i], dst_domain[i]), 3)

struct CoalescedPattern          # src/fft.jl:66-75 (one-device fields)
    src_idx::Int
    dst_idx::Int
    overlap::NTuple{3,UnitRange{Int}}
    src_indices::NTuple{3,UnitRange{Int}}
    dst_indices::NTuple{3,UnitRange{Int}}
    buffer_offset::Int
    buffer_size::Int
end

struct BlockedWorkspace
    local_patterns::Vector{CoalescedPattern}
end
cleanup_workspace!(::BlockedWorkspace) = lock(() -> empty!(PLAN_CACHE), PLAN_LOCK)

function create_workspace(src::BlockedArray{T}, dst::BlockedArray{T}; kwargs...) where {T}
    src.size == dst.size || throw(ArgumentError("create_workspace: arrays must share their size"))
    pats = CoalescedPattern[]
    for (dst_idx, dd) in enumerate(dst.subdomains), (src_idx, sd) in enumerate(src.subdomains)
        ov = compute_overlap(sd, dd)
        any(isempty, ov) && continue
        si = ntuple(i -> ov[i] .- first(sd[i]) .+ 1, 3)
        di = ntuple(i -> ov[i] .- first(dd[i]) .+ 1, 3)
        push!(pats, CoalescedPattern(src_idx, dst_idx, ov, si, di, 0, prod(length.(ov))))
    end
    return BlockedWorkspace(pats)
end

function redistribute!(dst::BlockedArray, src::BlockedArray, ws::BlockedWorkspace)
    for p in ws.local_patterns
        box_copy!(dst.chunks[p.dst_idx], first.(p.dst_indices), src.chunks[p.src_idx], first.(p.src_indices),
                  length.(p.overlap))
    end
    return dst
end

function apply_fft!(chunk::ROCArray{T}, transform, dims) where {T}
    d = dims isa Tuple ? dims : (dims,)
    if transform isa R2R
        ks = length(transform.kinds) == length(d) ? transform.kinds : ntuple(_ -> transform.kinds[1], length(d))
        p = cached((size(chunk), T, :r2r, d, ks, Pencil()),
                   () -> Plan(size(chunk), DFFT_R2R, -1, prec_code(T), Pencil(), d; kinds=ks))
    elseif transform isa Union{FFT,FFT!}
        p = plan_fft(chunk, d)
    elseif transform isa Union{IFFT,IFFT!}
        p = plan_ifft(chunk, d)
    else
        throw(ArgumentError("Unknown transform type"))
    end
    return execute!(p, chunk, chunk)
end

stage!(D::BlockedArray, transform, dims) = (foreach(ch -> apply_fft!(ch, transform, dims), D.chunks); D)

function fft!(C::BlockedArray, A::BlockedArray, B::BlockedArray, ws_AB::BlockedWorkspace, ws_BC::BlockedWorkspace,
              transforms, dims)
    stage!(A, transforms[1], dims[1]); redistribute!(B, A, ws_AB)
    stage!(B, transforms[2], dims[2]); redistribute!(C, B, ws_BC)
    return stage!(C, transforms[3], dims[3])
end
function ifft!(C::BlockedArray, A::BlockedArray, B::BlockedArray, ws_AB::BlockedWorkspace, ws_BC::BlockedWorkspace,
               transforms, dims)
    stage!(A, transforms[3], dims[3]); redistribute!(B, A, ws_AB)
    stage!(B, transforms[2], dims[2]); redistribute!(C, B, ws_BC)
    return stage!(C, transforms[1], dims[1])
end
function fft!(B::BlockedArray, A::BlockedArray, ws_AB::BlockedWorkspace, transforms, dims, ::Slab)
    stage!(A, transforms[1], (dims[1], dims[2])); redistribute!(B, A, ws_AB)
    return stage!(B, transforms[3], dims[3])
end
function ifft!(A::BlockedArray, B::BlockedArray, ws_BA::BlockedWorkspace, transforms, dims, ::Slab)
    stage!(B, transforms[3], dims[3]); redistribute!(A, B, ws_BA)
    return stage!(A, transforms[1], (dims[1], dims[2]))
end

# ---- batched transforms (dfft_batch_*; no reference counterpart) ----------
# Many independent 3D transforms per call: one grouped launch per axis-pass
# level (DFFT_FLAG_BATCH_GROUPED), the same tiles in one cooperative launch
# with a grid barrier between levels (DFFT_FLAG_BATCH_PERSISTENT), or each
# transform's own plan on four streams forked from the caller's
# (DFFT_FLAG_BATCH_LANES); by default the schedule cache's measured pick
# (DFFT_FLAG_MEASURE times the three).  DFFT_FLAG_GRAPH replays the launches
# as one hipGraph.
const DFFT_FLAG_GRAPH = UInt32(0x10)
const DFFT_FLAG_BATCH_GROUPED = UInt32(0x80)
const DFFT_FLAG_BATCH_LANES = UInt32(0x100)
const DFFT_FLAG_MEASURE = UInt32(0x200)
const DFFT_FLAG_BATCH_PERSISTENT = UInt32(0x1000)

struct BatchDesc            # dfft_batch_desc, field for field
    n::NTuple{3,Int64}
    kind::Cint
    sign::Cint
    precision::Cint
    flags::UInt32
end
BatchDesc(sz::Dims{3}, kind, sign, ::Type{T}) where {T} =
    BatchDesc(Int64.(sz), Cint(kind), Cint(sign), Cint(prec_code(T)), UInt32(0))

mutable struct Batch
    handle::Ptr{Cvoid}
    count::Int
end

function Batch(descs::Vector{BatchDesc}; flags=UInt32(0))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:dfft_batch_create, libdfft), Cint, (Ptr{Ptr{Cvoid}}, Cint, Ptr{BatchDesc}, UInt32),
                h, length(descs), descs, flags))
    b = Batch(h[], length(descs))
    finalizer(b) do q
        q.handle != C_NULL && ccall((:dfft_batch_destroy, libdfft), Cint, (Ptr{Cvoid},), q.handle)
        q.handle = C_NULL
    end
    return b
end

function execute!(b::Batch, outs::Vector{<:ROCArray}, ins::Vector{<:ROCArray})
    length(ins) == length(outs) == b.count || throw(ArgumentError("batch of $(b.count) transforms"))
    pin = Ptr{Cvoid}[pointer(x) for x in ins]
    pout = Ptr{Cvoid}[pointer(y) for y in outs]
    check(ccall((:dfft_batch_execute, libdfft), Cint, (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}),
                b.handle, pin, pout, stream_ptr()))
    return outs
end

# ---- serving loop (dfft_server_*; no reference counterpart) ---------------
# Requests are queued without waiting for the device, admitted under an HBM
# budget as slots and memory free up, and overlapped on several streams;
# plans and device memory stay resident.  The input of a request must hold
# its data when submitted (the server's streams do not wait for the task's
# stream: synchronize it first).
struct ServerConfig         # dfft_server_config
    hbm_budget::Csize_t
    max_inflight::Cint
    lanes::Cint
    flags::UInt32
    reserved::Cint
end

struct Request              # dfft_request
    n::NTuple{3,Int64}
    kind::Cint
    sign::Cint
    precision::Cint
    decomp::Cint
    flags::UInt32
    req_flags::UInt32
    seed::UInt64
    d_in::Ptr{Cvoid}
    d_out::Ptr{Cvoid}
    t_arrival::Cdouble
end

struct RequestInfo          # dfft_request_info
    state::Cint
    status::Cint
    lane::Cint
    reserved::Cint
    t_arrival::Cdouble
    t_admit::Cdouble
    t_done::Cdouble
    d_result::Ptr{Cvoid}
    result_bytes::Csize_t
    algorithmic_bytes::Cdouble
end

mutable struct Server
    handle::Ptr{Cvoid}
end

function Server(; hbm_budget=200 * 10^9, max_inflight=32, lanes=4)
    cfg = Ref(ServerConfig(Csize_t(hbm_budget), Cint(max_inflight), Cint(lanes), UInt32(0), Cint(0)))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:dfft_server_create, libdfft), Cint, (Ptr{Ptr{Cvoid}}, Ptr{ServerConfig}), h, cfg))
    s = Server(h[])
    finalizer(s) do q
        q.handle != C_NULL && ccall((:dfft_server_destroy, libdfft), Cint, (Ptr{Cvoid},), q.handle)
        q.handle = C_NULL
    end
    return s
end

const DFFT_REQ_DISCARD = UInt32(0x1)

# fft (complex input) or rfft (real input) of a 3D array, or ifft with sign=+1;
# with `out`, the result is written there, else it stays in the server until
# release (discard=true: nobody reads it, and its server buffer is freed when
# the request completes).  Returns the ticket.
function submit(s::Server, A::ROCArray{T,3}; out=nothing, sign=-1, decomp=Pencil(), discard=false) where {T}
    kind = T <: Real ? DFFT_R2C : DFFT_C2C
    req = Ref(Request(Int64.(size(A)), Cint(kind), Cint(sign), Cint(prec_code(T)), Cint(decomp_code(decomp)),
                      UInt32(0), discard ? DFFT_REQ_DISCARD : UInt32(0), UInt64(0), pointer(A),
                      out === nothing ? C_NULL : pointer(out), 0.0))
    t = Ref{Int64}(0)
    check(ccall((:dfft_server_submit, libdfft), Cint, (Ptr{Cvoid}, Ptr{Request}, Ptr{Int64}), s.handle, req, t))
    return t[]
end

# true once the request finished (timeout < 0: wait without limit)
function wait_ticket(s::Server, ticket::Int64; timeout=-1.0)
    st = ccall((:dfft_server_wait, libdfft), Cint, (Ptr{Cvoid}, Int64, Cdouble), s.handle, ticket, timeout)
    st == 7 && return false   # DFFT_ERR_NOT_READY
    check(st)
    info = Ref{RequestInfo}()
    check(ccall((:dfft_server_info, libdfft), Cint, (Ptr{Cvoid}, Int64, Ptr{RequestInfo}), s.handle, ticket, info))
    check(info[].status)
    return true
end

# copy a finished request's result into `out` (ordered on the task's stream)
function result!(out::ROCArray, s::Server, ticket::Int64)
    check(ccall((:dfft_server_copy_result, libdfft), Cint, (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Csize_t, Ptr{Cvoid}),
                s.handle, ticket, pointer(out), sizeof(out), stream_ptr()))
    return out
end

release(s::Server, ticket::Int64) =
    check(ccall((:dfft_server_release, libdfft), Cint, (Ptr{Cvoid}, Int64), s.handle, ticket))

end # module
