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
struct Workspace
    src_size::Dims
    dst_size::Dims
end
create_workspace(src::ROCArray, dst::ROCArray; kwargs...) = Workspace(size(src), size(dst))
cleanup_workspace!(::Workspace) = lock(() -> empty!(PLAN_CACHE), PLAN_LOCK)

stage_inverse(transforms) = first(transforms) isa Union{IFFT,IFFT!}

function fft!(C::ROCArray, A::ROCArray, B::ROCArray, ws_AB::Workspace, ws_BC::Workspace, transforms, dims)
    return execute!(plan_fft(A, dims; decomp=Pencil()), C, A)
end
function fft!(B::ROCArray, A::ROCArray, ws::Workspace, transforms, dims, ::Slab)
    return execute!(plan_fft(A, dims; decomp=Slab()), B, A)
end
function ifft!(C::ROCArray, A::ROCArray, B::ROCArray, ws_AB::Workspace, ws_BC::Workspace, transforms, dims)
    return execute!(plan_ifft(A, dims; decomp=Pencil()), C, A)
end
function ifft!(A::ROCArray, B::ROCArray, ws::Workspace, transforms, dims, ::Slab)
    return execute!(plan_ifft(B, dims; decomp=Slab()), A, B)
end

end # module
