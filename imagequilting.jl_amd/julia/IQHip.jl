# IQHip.jl -- ccall binding of libiqhip (include/iqhip.h) for ImageQuilting.jl.
#
# `iqsim_hip` keeps `iqsim`'s signature and return convention
# (src/iqsim.jl:50-63, 314) and reuses ImageQuilting's own host preparation
# (imagepreproc, finddisabled, findskipped, genpath; src/utils.jl:69-204);
# the tile loop (src/iqsim.jl:163-285) runs on the GPU through iq_simulate.
# The per-function wrappers replace imfilter_kernel (src/imfilter.jl:28),
# fastdistance (src/utils.jl:5), relaxation, taumodel and graphcut.
#
# Not exercised in this repository (no Julia toolchain here); the same ABI is
# tested from Python through ctypes.
module IQHip

using ImageQuilting
using Random

const libiqhip = get(ENV, "IQHIP_LIB", joinpath(@__DIR__, "..", "lib", "libiqhip.so"))

const IQ_F32 = Cint(0)
const IQ_F64 = Cint(1)

struct PlanDesc
  ti::Ptr{Cvoid}
  ti_dtype::Cint
  tisize::NTuple{3,Int64}
  tilesize::NTuple{3,Int64}
  ovlsize::NTuple{3,Int64}
  spacing::NTuple{3,Int64}
  ntiles::NTuple{3,Int64}
  padsize::NTuple{3,Int64}
  disabled::Ptr{UInt8}
  path::Ptr{Int64}
  npath::Int64
  skipped::Ptr{UInt8}
  hardgrid::Ptr{Float64}
  naux::Cint
  aux::Ptr{Ptr{Float64}}
  auxti::Ptr{Ptr{Float64}}
  tol::Float64
  nreal::Int64
end

# extents padded to three axes: a missing trailing axis is one voxel thick
pad3(t::NTuple{N,Int}) where {N} = ntuple(i -> i <= N ? Int64(t[i]) : Int64(1), 3)
# overlaps padded to three axes: a missing trailing axis has no overlap
# (spacing = tilesize - ovlsize = 1 and padsize = ntiles * spacing + ovlsize = 1 there)
pad3ovl(t::NTuple{N,Int}) where {N} = ntuple(i -> i <= N ? Int64(t[i]) : Int64(0), 3)

function check(rc, ctx=C_NULL)
  rc == 0 && return nothing
  msg = unsafe_string(ccall((:iq_last_error, libiqhip), Cstring, (Ptr{Cvoid},), ctx))
  error("libiqhip error $rc: $msg")
end

mutable struct Context
  handle::Ptr{Cvoid}
  function Context(device::Integer=0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:iq_ctx_create, libiqhip), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h))
    ctx = new(h[])
    finalizer(c -> ccall((:iq_ctx_destroy, libiqhip), Cint, (Ptr{Cvoid},), c.handle), ctx)
    ctx
  end
end

const DEFAULT = Ref{Union{Nothing,Context}}(nothing)
context() = (DEFAULT[] === nothing && (DEFAULT[] = Context(0)); DEFAULT[])

# imfilter_kernel replacement (src/imfilter.jl:5-28)
function imfilter_hip(img::AbstractArray{<:Real,N}, krn::AbstractArray{<:Real,N}; ctx=context()) where {N}
  A = Array{Float64}(img)
  K = Array{Float64}(krn)
  out = Array{Float64}(undef, size(A) .- size(K) .+ 1)
  check(ccall((:iq_imfilter, libiqhip), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Ref{NTuple{3,Int64}}, Ptr{Float64}, Ref{NTuple{3,Int64}}, Ptr{Float64}),
              ctx.handle, A, pad3(size(A)), K, pad3(size(K)), out), ctx.handle)
  out
end

# fastdistance replacement (src/utils.jl:5-13)
function fastdistance_hip(img::AbstractArray{T,N}, kern; weights=fill(true, size(kern)), ctx=context()) where {T,N}
  A = T === Float32 ? Array{Float32}(img) : Array{Float64}(img)
  dt = T === Float32 ? IQ_F32 : IQ_F64
  K = Array{Float64}(kern)
  W = Array{UInt8}(weights .!= 0)
  out = Array{Float64}(undef, size(A) .- size(K) .+ 1)
  check(ccall((:iq_fastdistance, libiqhip), Cint,
              (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{NTuple{3,Int64}}, Ptr{Float64}, Ptr{UInt8}, Ref{NTuple{3,Int64}},
               Ptr{Float64}), ctx.handle, A, dt, pad3(size(A)), K, W, pad3(size(K)), out), ctx.handle)
  out
end

# graphcut replacement (src/graphcut.jl:5-84); dim is 1-based as in Julia
function graphcut_hip(A::AbstractArray{T,N}, B::AbstractArray{T,N}, dim::Integer; ctx=context()) where {T,N}
  dt = T === Float32 ? IQ_F32 : IQ_F64
  a = T === Float32 ? Array{Float32}(A) : Array{Float64}(A)
  b = T === Float32 ? Array{Float32}(B) : Array{Float64}(B)
  m = Array{UInt8}(undef, size(a))
  check(ccall((:iq_graphcut, libiqhip), Cint,
              (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ref{NTuple{3,Int64}}, Cint, Ptr{UInt8}),
              ctx.handle, a, b, dt, pad3(size(a)), dim - 1, m), ctx.handle)
  BitArray(m .!= 0)
end

# raw element types of include/iqhip.h (iq_raw_dtype); others go through Float64
const RAWCODE = Dict(Float32 => 0, Float64 => 1, UInt8 => 2, Int8 => 3, UInt16 => 4, Int16 => 5, UInt32 => 6,
                     Int32 => 7, UInt64 => 8, Int64 => 9, Bool => 2)

function rawarray(a::AbstractArray)
  T = nonmissingtype(eltype(a))
  S = T === Bool ? UInt8 : (haskey(RAWCODE, T) ? T : Float64)
  data = Array{S}(undef, size(a))
  miss = zeros(UInt8, size(a))
  for (k, v) in enumerate(a)
    if ismissing(v)
      miss[k] = 1
      data[k] = zero(S)
    else
      data[k] = S(v)
    end
  end
  data, Cint(RAWCODE[S]), miss
end

"prepare (src/utils.jl:96-101) in the library."
function prepare_hip(a::AbstractArray)
  data, code, miss = rawarray(a)
  out = Array{code == 0 ? Float32 : Float64}(undef, size(data))
  check(ccall((:iq_prepare, libiqhip), Cint, (Ptr{Cvoid}, Cint, Ptr{UInt8}, Ref{NTuple{3,Int64}}, Ptr{Cvoid}),
              data, code, miss, pad3(size(data)), out))
  out
end

"Soft grid as the tile loop reads it (src/utils.jl:74-80), padsize Float64."
function prepare_soft_hip(aux::AbstractArray, padsize::Tuple)
  data, code, miss = rawarray(aux)
  out = Array{Float64}(undef, padsize)
  check(ccall((:iq_prepare_soft, libiqhip), Cint,
              (Ptr{Cvoid}, Cint, Ptr{UInt8}, Ref{NTuple{3,Int64}}, Ref{NTuple{3,Int64}}, Ptr{Float64}),
              data, code, miss, pad3(size(data)), pad3(padsize), out))
  out
end

"finddisabled (src/utils.jl:115-129) in the library."
function finddisabled_hip(trainimg::AbstractArray, tilesize::Tuple)
  data, code, miss = rawarray(trainimg)
  out = Array{UInt8}(undef, size(trainimg) .- tilesize .+ 1)
  check(ccall((:iq_finddisabled, libiqhip), Cint,
              (Ptr{Cvoid}, Cint, Ptr{UInt8}, Ref{NTuple{3,Int64}}, Ref{NTuple{3,Int64}}, Ptr{UInt8}),
              data, code, miss, pad3(size(data)), pad3(tilesize), out))
  out
end

"""
    iqsim_hip(trainimg, tilesize, simsize=size(trainimg); kwargs...)

Same keywords, validation and return value as `ImageQuilting.iqsim`; the tile
loop runs on the GPU.  Draws are taken from `rng` and handed to the device.
"""
function iqsim_hip(trainimg::AbstractArray{T,N}, tilesize::Dims{N}, simsize::Dims{N}=size(trainimg);
                   overlap::NTuple{N,<:Real}=ntuple(i -> 1 / 6, N), soft::AbstractVector=[], hard::Dict=Dict(),
                   tol::Real=0.1, path::Symbol=:raster, nreal::Integer=1, debug::Bool=false,
                   showprogress::Bool=false, rng=Random.default_rng(), ctx=context()) where {T,N}
  # sanity checks and geometry exactly as src/iqsim.jl:69-127
  @assert all(0 .< tilesize .≤ size(trainimg)) "invalid tile size"
  @assert all(simsize .≥ tilesize) "invalid grid size"
  @assert all(0 .< overlap .< 1) "overlaps must be in range (0,1)"
  @assert 0 < tol ≤ 1 "tolerance must be in range (0,1]"
  @assert path ∈ [:raster, :dilation, :random] "invalid simulation path"
  @assert nreal > 0 "invalid number of realizations"
  ovlsize = @. ceil(Int, overlap * tilesize)
  spacing = @. tilesize - ovlsize
  ntiles = @. ceil(Int, simsize / max(spacing, 1))
  padsize = @. ntiles * (tilesize - ovlsize) + ovlsize
  TIsize = size(trainimg)
  distsize = TIsize .- tilesize .+ 1
  ovlvol = prod(padsize) - prod(@. padsize - (ntiles - 1) * ovlsize)
  geo = (ntiles=ntiles, tilesize=tilesize, ovlsize=ovlsize, spacing=spacing, TIsize=TIsize, simsize=simsize,
         padsize=padsize, distsize=distsize)
  # type conversion, soft padding and disabled windows inside the library
  # (src/utils.jl:69-129): raw arrays in
  TI = prepare_hip(trainimg)
  SOFT = map(soft) do (aux, auxTI)
    prepare_soft_hip(aux, Tuple(padsize)), Array{Float64}(prepare_hip(auxTI))
  end
  disabled = finddisabled_hip(trainimg, tilesize)
  # tile skipping and paths reused from ImageQuilting (src/utils.jl:131-204)
  skipped, datainds = ImageQuilting.findskipped(hard, geo)
  simpath = Int64.(ImageQuilting.genpath(rng, ntiles, path, datainds)) .- 1
  skp = zeros(UInt8, prod(ntiles))
  for s in skipped
    skp[s] = 1
  end
  hardgrid = fill(NaN, padsize)
  for (coord, val) in hard
    hardgrid[coord] = val
  end
  tidt = eltype(TI) === Float32 ? IQ_F32 : IQ_F64
  tiarr = eltype(TI) === Float32 ? Array{Float32}(TI) : Array{Float64}(TI)
  auxp = [pointer(a) for (a, _) in SOFT]
  auxtip = [pointer(b) for (_, b) in SOFT]
  uniforms = rand(rng, length(simpath), nreal)  # column r holds realization r's draws
  out = Array{eltype(tiarr)}(undef, padsize..., nreal)
  chosen = Array{Int64}(undef, length(simpath), nreal)
  cuts = debug ? Array{Float64}(undef, padsize..., nreal) : Float64[]
  GC.@preserve tiarr SOFT disabled simpath skp hardgrid auxp auxtip uniforms out chosen cuts begin
    desc = PlanDesc(pointer(tiarr), tidt, pad3(TIsize), pad3(tilesize), pad3ovl(Tuple(ovlsize)), pad3(Tuple(spacing)),
                    pad3(Tuple(ntiles)), pad3(Tuple(padsize)), pointer(disabled), pointer(simpath), length(simpath),
                    pointer(skp), isempty(hard) ? Ptr{Float64}(C_NULL) : pointer(hardgrid), Cint(length(SOFT)),
                    isempty(SOFT) ? Ptr{Ptr{Float64}}(C_NULL) : pointer(auxp),
                    isempty(SOFT) ? Ptr{Ptr{Float64}}(C_NULL) : pointer(auxtip), Float64(tol), Int64(nreal))
    check(ccall((:iq_simulate, libiqhip), Cint,
                (Ptr{Cvoid}, Ref{PlanDesc}, Ptr{Float64}, Ptr{Int64}, Ptr{Cvoid}, Ptr{Int64}, Ptr{Float64}),
                ctx.handle, desc, uniforms, C_NULL, out, chosen, debug ? pointer(cuts) : C_NULL), ctx.handle)
  end
  # hard-data overwrite, crop and typing (src/iqsim.jl:287-308)
  R = T <: AbstractFloat ? T : Union{Missing,nonmissingtype(T)}
  reals = Vector{Array{R,N}}()
  boundarycuts = Vector{Array{Float64,N}}()
  voxs = Float64[]
  for r in 1:nreal
    simgrid = copy(selectdim(out, N + 1, r))
    if debug
      cutgrid = copy(selectdim(cuts, N + 1, r))
      push!(voxs, sum(cutgrid) / ovlvol)
    end
    for (coord, val) in hard
      simgrid[coord] = val
      debug && isnan(val) && (cutgrid[coord] = val)
    end
    push!(reals, ImageQuilting.imagepostproc(R, simgrid[CartesianIndices(simsize)]))
    debug && push!(boundarycuts, cutgrid[CartesianIndices(simsize)])
  end
  debug ? (reals, boundarycuts, voxs) : reals
end

export iqsim_hip, imfilter_hip, fastdistance_hip, graphcut_hip

end # module
