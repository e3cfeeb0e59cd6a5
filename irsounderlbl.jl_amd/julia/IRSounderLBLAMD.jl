"""
    IRSounderLBLAMD

Julia binding of libirs.so (include/irs.h): `forward_model` and
`analytic_jacobian` with the reference signatures (src/Sensor/sounder.jl:233-252,
src/Estimation/vmr_jacobian.jl:152-173), each a single `ccall` per batch into
hand-written gfx950 kernels.  The reference's own types (`AtmosphericProfile`,
`HITRANLinelist`, `Sounder`, `ViewingGeometry`, `StateVectorSpec`, `Jacobian`,
`WavenumberGrid`) are used unchanged; include this file after `using IRSounderLBL`.

Line mixing: `line_mixing = VPYLineMixing(relmat)` or `VPWLineMixing(relmat; ...)`
(the reference's types) uploads the relaxation matrix once and evaluates Y(T),
the VP_W eigenmodes and their sums per layer in libirs; `line_mixing =
Y::Vector{Float64}` (atm^-1, one per line of the CO2 list, in list order) applies
a fixed first-order coefficient per line instead.  Continuum (`apply_continuum`, `continua`)
is computed on the device from the bundled MT-CKD tables; the HITRAN CIA pairs
add zero until a table is loaded with `load_cia!`, as in the reference when the
data files are absent.
"""
module IRSounderLBLAMD

using IRSounderLBL: AtmosphericProfile, HITRANLine, HITRANLinelist, GasSpecies, Sounder, IASIInstrument,
                    ViewingGeometry, nadir_geometry, WavenumberGrid, wavenumber_grid,
                    StateVectorSpec, Jacobian, H2O, CO2
using IRSounderLBL: VPYLineMixing, VPWLineMixing, HITRANRelmatData
import IRSounderLBL

const libirs = get(ENV, "IRS_LIBRARY",
                   joinpath(@__DIR__, "..", "csrc", "libirs.so"))

# ---- C structs (layout of include/irs.h) ------------------------------------
struct CLinelist
    n::Int64
    mol::Ptr{Int8}; iso::Ptr{Int8}
    nu::Ptr{Float64}; S::Ptr{Float64}
    gamma_air::Ptr{Float32}; gamma_self::Ptr{Float32}
    elower::Ptr{Float64}; n_air::Ptr{Float32}; delta_air::Ptr{Float32}
    y_lm::Ptr{Float64}
end

struct CParLines
    mol::Ptr{Int8}; iso::Ptr{Int8}
    nu::Ptr{Float64}; S::Ptr{Float64}; A::Ptr{Float64}; elower::Ptr{Float64}
    gamma_air::Ptr{Float32}; gamma_self::Ptr{Float32}; n_air::Ptr{Float32}; delta_air::Ptr{Float32}
end

struct CSounder
    nu_min::Float64; nu_max::Float64; dnu::Float64
    opd_max::Float64; fwhm_gauss::Float64
    n_channels::Int64
end

struct CProfile
    n_levels::Cint
    p_hpa::Ptr{Float64}; T::Ptr{Float64}
    n_species::Cint
    species::Ptr{Cint}; linelist::Ptr{Int32}; vmr::Ptr{Float64}
    T_sfc::Float64; eps_sfc::Float64; mu::Float64
end

mutable struct COptions
    internal_dnu::Float64; high_res_factor::Cint; cutoff::Float64; lm_cutoff::Float64
    dptmn::Float64; x_far::Float64; with_ils::Cint; apodization::Cint; line_mixing::Cint
    T_method::Cint; source_function::Cint; continua::Cint
    relmat::Int32; lm_species::Int32; relmat_band_mode::Ptr{UInt8}; keep_threshold::Float64
end

struct CRelmat
    n_bands::Int32
    li::Ptr{Int8}; lf::Ptr{Int8}; isot::Ptr{Int8}; molecule::Ptr{Int8}
    line_off::Ptr{Int64}
    Ji::Ptr{Int16}; branch::Ptr{Int8}
    nu::Ptr{Float64}; gV_air::Ptr{Float64}; E_lower::Ptr{Float64}; Dipo0::Ptr{Float64}
    PopuT0::Ptr{Float64}; DipoT::Ptr{Float64}
    n_air::Ptr{Float32}; shift::Ptr{Float32}
    n_tables::Int32
    table_lli::Ptr{Int8}; table_llf::Ptr{Int8}
    entry_off::Ptr{Int64}
    entry_sub::Ptr{Int8}
    entry_ji::Ptr{Int32}; entry_jip::Ptr{Int32}
    entry_w0::Ptr{Float64}; entry_b0::Ptr{Float64}
end

const SUBTABLES = (:pp, :pq, :pr, :qp, :qq, :qr, :rp, :rq, :rr)

struct CStateSpec
    include_temperature::Cint; n_retrieved::Cint
    retrieved::Ptr{Cint}; column_scale::Ptr{Cint}
    include_tsfc::Cint; include_emissivity::Cint; log_vmr::Cint; vmr_coupling::Cint
    observable::Cint
end

const APOD = Dict(:gaussian => 0, :norton_beer_weak => 1, :norton_beer_medium => 2,
                  :norton_beer_strong => 3)
const CONTINUA = Dict(:h2o => 1, :co2 => 2, :co2_cia => 4, :n2 => 8, :o2 => 16)
const CIA_PAIR = Dict(:co2_cia => 0, :n2 => 1, :o2 => 2)

# ---- context and line-list cache --------------------------------------------
mutable struct Context
    h::Ptr{Cvoid}
    lists::IdDict{Any, Int32}          # HITRANLinelist (+ Y) => device handle
    keep::Vector{Any}
end

function check(ctx::Context, rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:irs_last_error, libirs), Cstring, (Ptr{Cvoid},), ctx.h))
    error(msg)                          # same ErrorException text as the reference check
end

const _CTX = Ref{Union{Nothing, Context}}(nothing)

function context(device::Integer = parse(Int, get(ENV, "IRS_DEVICE", "0")))
    _CTX[] === nothing || return _CTX[]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:irs_create, libirs), Cint, (Cint, Ref{Ptr{Cvoid}}), device, h)
    rc == 0 || error("irs_create failed ($rc): no usable gfx950 device")
    ctx = Context(h[], IdDict{Any, Int32}(), Any[])
    finalizer(c -> ccall((:irs_destroy, libirs), Cvoid, (Ptr{Cvoid},), c.h), ctx)
    _CTX[] = ctx
end

function upload!(ctx::Context, ll::HITRANLinelist, y::Union{Nothing, Vector{Float64}})
    key = y === nothing ? ll : (ll, y)
    haskey(ctx.lists, key) && return ctx.lists[key]
    L = ll.lines
    mol = Int8[l.mol_id for l in L];        iso = Int8[l.iso_id for l in L]
    nu = Float64[l.wavenumber for l in L];  S = Float64[l.intensity for l in L]
    ga = Float32[l.air_broad for l in L];   gs = Float32[l.self_broad for l in L]
    el = Float64[l.lower_energy for l in L]; na = Float32[l.temp_depend for l in L]
    da = Float32[l.pressure_shift for l in L]
    y === nothing || length(y) == length(L) || error("line_mixing Y must have one entry per CO2 line")
    h = Ref{Int32}(-1)
    GC.@preserve mol iso nu S ga gs el na da y begin
        c = CLinelist(length(L), pointer(mol), pointer(iso), pointer(nu), pointer(S), pointer(ga),
                      pointer(gs), pointer(el), pointer(na), pointer(da),
                      y === nothing ? Ptr{Float64}(C_NULL) : pointer(y))
        check(ctx, ccall((:irs_linelist_upload, libirs), Cint,
                         (Ptr{Cvoid}, Ref{CLinelist}, Ref{Int32}), ctx.h, c, h))
    end
    ctx.lists[key] = h[]
end

"""
    load_linelist(parpath; ν_min=nothing, ν_max=nothing, cache=true, rebuild=false) -> HITRANLinelist

`load_linelist` (src/HITRAN/cache.jl:153-175) with the records parsed on the
device (`irs_par_load`): the whole .par parsed once, cached as a
structure-of-arrays file in the reference's `linelist_cache_dir()` (keyed by
path, size, mtime and the cache version), the window applied in memory.
"""
function load_linelist(parpath::AbstractString; ν_min::Union{Float64,Nothing} = nothing,
                       ν_max::Union{Float64,Nothing} = nothing, cache::Bool = true, rebuild::Bool = false,
                       ctx::Context = context())
    isfile(parpath) || error("File not found: $parpath")
    cap = filesize(parpath) ÷ 67 + 1
    mol = Vector{Int8}(undef, cap); iso = Vector{Int8}(undef, cap)
    nu = Vector{Float64}(undef, cap); S = Vector{Float64}(undef, cap); A = Vector{Float64}(undef, cap)
    el = Vector{Float64}(undef, cap); ga = Vector{Float32}(undef, cap); gs = Vector{Float32}(undef, cap)
    na = Vector{Float32}(undef, cap); da = Vector{Float32}(undef, cap)
    lo = something(ν_min, -Inf); hi = something(ν_max, Inf)
    dir = cache ? IRSounderLBL.linelist_cache_dir() : ""
    n = Ref{Int64}(0)
    GC.@preserve mol iso nu S A el ga gs na da dir begin
        out = CParLines(pointer(mol), pointer(iso), pointer(nu), pointer(S), pointer(A), pointer(el),
                        pointer(ga), pointer(gs), pointer(na), pointer(da))
        check(ctx, ccall((:irs_par_load, libirs), Cint,
                         (Ptr{Cvoid}, Cstring, Cstring, Cint, Float64, Float64, Int64, Ref{CParLines}, Ref{Int64}),
                         ctx.h, parpath, dir, rebuild, lo, hi, cap, out, n))
    end
    lines = [HITRANLine(mol[i], iso[i], nu[i], S[i], A[i], ga[i], gs[i], el[i], na[i], da[i]) for i in 1:n[]]
    isempty(lines) && return HITRANLinelist(lines, Set{Int}(), lo, hi)
    return HITRANLinelist(lines)
end

function continua_mask(apply_continuum::Bool, continua)
    apply_continuum || return Cint(0)
    m = 0
    for c in continua
        haskey(CONTINUA, c) || error("unknown continuum :$c; expected one of $(collect(keys(CONTINUA)))")
        m |= CONTINUA[c]
    end
    return Cint(m)
end

function options(; internal_dnu, high_res_factor, cutoff, lm_cutoff, dptmn, with_ils,
                 apodization, line_mixing, T_method, source_function, x_far = 122.0,
                 apply_continuum = false, continua = ())
    o = COptions(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, C_NULL, 0)
    ccall((:irs_default_options, libirs), Cvoid, (Ref{COptions},), o)
    o.internal_dnu = internal_dnu === nothing ? 0.0 : internal_dnu
    o.high_res_factor = high_res_factor === nothing ? 0 : high_res_factor
    o.cutoff = cutoff; o.lm_cutoff = lm_cutoff; o.dptmn = dptmn; o.x_far = x_far
    o.with_ils = with_ils
    haskey(APOD, apodization) ||
        error("Unknown apodization :$apodization; expected :gaussian or a Norton-Beer style")
    o.apodization = APOD[apodization]
    o.line_mixing = line_mixing === nothing ? 0 :
                    (line_mixing isa Union{VPYLineMixing, VPWLineMixing} ? 2 : 1)
    T_method in (:logp_at_pcg, :mass_weighted) ||
        error("T_method must be :logp_at_pcg or :mass_weighted, got :$T_method")
    o.T_method = T_method === :logp_at_pcg ? 0 : 1
    source_function in (:toon, :cim) || error("source_function must be :toon or :cim, got :$source_function")
    o.source_function = source_function === :toon ? 0 : 1
    o.continua = continua_mask(apply_continuum, continua)
    return o
end

"""
    load_cia!(pair, path; ctx = context())

Load a HITRAN `.cia` file for `pair` (`:co2_cia`, `:n2` or `:o2`, the `continua`
symbols of the pair) into the device table (src/CrossSections/continuum.jl:106-148).
"""
function load_cia!(pair::Symbol, path::AbstractString; ctx::Context = context())
    haskey(CIA_PAIR, pair) || error("CIA pair must be :co2_cia, :n2 or :o2, got :$pair")
    check(ctx, ccall((:irs_cia_load_file, libirs), Cint, (Ptr{Cvoid}, Cint, Cstring),
                     ctx.h, CIA_PAIR[pair], path))
end

csounder(s::Sounder) = CSounder(s.ν_min, s.ν_max, s.Δν, s.opd_max, s.fwhm_gauss, s.n_channels)

# ---- relaxation matrices (line_mixing.jl:35-117) ------------------------------------
function upload_relmat!(ctx::Context, data::HITRANRelmatData)
    haskey(ctx.lists, data) && return ctx.lists[data]
    B = data.bands
    li = Int8[b.li for b in B]; lf = Int8[b.lf for b in B]; isot = Int8[b.isot for b in B]
    mol = Int8[b.molecule for b in B]
    off = Int64[0; cumsum([length(b.lines) for b in B])]
    L = reduce(vcat, [b.lines for b in B]; init = IRSounderLBL.RelmatLine[])
    Ji = Int16[l.Ji for l in L]; br = Int8[l.branch for l in L]
    nu = Float64[l.ν for l in L]; gv = Float64[l.gV_air for l in L]; el = Float64[l.E_lower for l in L]
    d0 = Float64[l.Dipo0 for l in L]; p0 = Float64[l.PopuT0 for l in L]; dt = Float64[l.DipoT for l in L]
    na = Float32[l.n_air for l in L]; sh = Float32[l.shift for l in L]
    keys_ = sort(collect(keys(data.wtfit)))
    lli = Int8[k[1] for k in keys_]; llf = Int8[k[2] for k in keys_]
    eoff = Int64[0]; esub = Int8[]; eji = Int32[]; ejip = Int32[]; ew0 = Float64[]; eb0 = Float64[]
    for k in keys_, (si, sub) in enumerate(SUBTABLES)
        for ((ji, jip), (w0, b0)) in getfield(data.wtfit[k], sub)
            push!(esub, si - 1); push!(eji, ji); push!(ejip, jip); push!(ew0, w0); push!(eb0, b0)
        end
        sub === :rr && push!(eoff, length(esub))
    end
    h = Ref{Int32}(-1)
    GC.@preserve li lf isot mol off Ji br nu gv el d0 p0 dt na sh lli llf eoff esub eji ejip ew0 eb0 begin
        c = CRelmat(length(B), pointer(li), pointer(lf), pointer(isot), pointer(mol), pointer(off),
                    pointer(Ji), pointer(br), pointer(nu), pointer(gv), pointer(el), pointer(d0), pointer(p0),
                    pointer(dt), pointer(na), pointer(sh), length(keys_), pointer(lli), pointer(llf),
                    pointer(eoff), pointer(esub), pointer(eji), pointer(ejip), pointer(ew0), pointer(eb0))
        check(ctx, ccall((:irs_relmat_upload, libirs), Cint, (Ptr{Cvoid}, Ref{CRelmat}, Ref{Int32}), ctx.h, c, h))
    end
    ctx.lists[data] = h[]
end

# Per-band treatment: the band filters of compute_lm_dispersive_correction and
# _vpw_perturbation (line_mixing.jl:1120-1154,1303-1349), isotopologues of the
# species' line list (line_mixing.jl:199-206).
function relmat_band_modes(lm, linelists)
    data = lm.data
    iso = haskey(linelists, data.species) ? IRSounderLBL._linelist_isotopes(linelists[data.species]) : nothing
    wl = lm isa VPWLineMixing ? lm.whitelist : Set{String}()
    m = zeros(UInt8, length(data.bands))
    for (i, b) in enumerate(data.bands)
        Int(b.li) > 8 && continue
        iso !== nothing && Int(b.isot) ∉ iso && continue
        lm.min_band_strength > 0.0 && IRSounderLBL._band_eff_strength(b) < lm.min_band_strength && continue
        haskey(data.wtfit, (Int8(min(b.li, b.lf)), Int8(max(b.li, b.lf)))) || continue
        m[i] = b.name in wl ? 2 : 1
    end
    return m
end

function relmat_options!(o::COptions, ctx::Context, line_mixing, linelists, keep)
    line_mixing isa Union{VPYLineMixing, VPWLineMixing} || return o
    modes = relmat_band_modes(line_mixing, linelists)
    push!(keep, modes)
    o.relmat = upload_relmat!(ctx, line_mixing.data)
    o.lm_species = Int(line_mixing.data.species)
    o.relmat_band_mode = isempty(modes) ? Ptr{UInt8}(C_NULL) : pointer(modes)
    return o
end

function cprofile(ctx, prof::AtmosphericProfile, linelists, line_mixing, T_sfc, ε_sfc, μ, keep)
    species = collect(keys(prof.vmr))
    sp = Cint[Int(s) for s in species]
    y = line_mixing isa AbstractVector ? line_mixing : nothing
    hs = Int32[haskey(linelists, s) && !isempty(linelists[s].lines) ?
               upload!(ctx, linelists[s], (s == CO2 && y !== nothing) ? y : nothing) :
               Int32(-1) for s in species]
    vmr = reduce(vcat, [Float64.(prof.vmr[s]) for s in species])
    p = Float64.(prof.pressure); T = Float64.(prof.temperature)
    append!(keep, (sp, hs, vmr, p, T))
    return CProfile(length(p), pointer(p), pointer(T), length(sp), pointer(sp), pointer(hs),
                    pointer(vmr), T_sfc === nothing ? 0.0 : T_sfc, ε_sfc, μ), species
end

"""
    forward_model(prof, linelists; sounder, geom, ...) -> (WavenumberGrid, R, BT)

src/Sensor/sounder.jl:233-360 on the GPU.  `prof` may also be a vector of
profiles (with `geom`, `T_sfc`, `ε_sfc` scalars or vectors): one call, one
device launch sequence; R and BT are then n_channels x n_profiles.
"""
function forward_model(prof, linelists::Dict{GasSpecies, HITRANLinelist};
                       sounder::Sounder = IASIInstrument(),
                       iasi::Union{Sounder, Nothing} = nothing, geom = nadir_geometry(),
                       T_sfc = nothing, ε_sfc = 1.0,
                       internal_dnu::Union{Float64, Nothing} = 0.001,
                       high_res_factor::Union{Int, Nothing} = nothing,
                       cutoff::Float64 = 25.0, lm_cutoff::Float64 = cutoff, dptmn::Float64 = 1e-6,
                       apply_continuum::Bool = true, continua = (:h2o, :co2, :co2_cia, :n2, :o2),
                       with_ils::Bool = true, apodization::Symbol = :gaussian,
                       line_mixing = nothing, T_method::Symbol = :logp_at_pcg,
                       source_function::Symbol = :cim, backend = nothing)
    snd = iasi === nothing ? sounder : iasi       # sounder.jl:254: `iasi=` wins when given
    profs = prof isa AbstractVector ? prof : [prof]
    nb = length(profs)
    pick(x, q) = x isa AbstractVector ? x[q] : x
    ctx = context()
    o = options(; internal_dnu, high_res_factor, cutoff, lm_cutoff, dptmn, with_ils, apodization,
                line_mixing, T_method, source_function, apply_continuum, continua)
    keep = Any[]
    relmat_options!(o, ctx, line_mixing, linelists, keep)
    cps = Vector{CProfile}(undef, nb)
    for q in 1:nb
        cps[q], _ = cprofile(ctx, profs[q], linelists, line_mixing, pick(T_sfc, q), pick(ε_sfc, q),
                             pick(geom, q).μ, keep)
    end
    grid = wavenumber_grid(snd.ν_min, snd.ν_max, snd.Δν)
    R = zeros(Float64, grid.n, nb); BT = zeros(Float64, grid.n, nb)
    cs = csounder(snd)
    GC.@preserve keep cps R BT begin
        check(ctx, ccall((:irs_forward, libirs), Cint,
                         (Ptr{Cvoid}, Ptr{CProfile}, Cint, Ref{CSounder}, Ref{COptions},
                          Ptr{Float64}, Ptr{Float64}),
                         ctx.h, cps, nb, cs, o, R, BT))
    end
    prof isa AbstractVector && return grid, R, BT
    return grid, R[:, 1], BT[:, 1]
end

"""
    analytic_jacobian(prof, linelists, spec; ...) -> Jacobian

src/Estimation/vmr_jacobian.jl:152-505 on the GPU (full line set, no DPTMIN).
FullProfile and ColumnScale blocks are assembled on the device; a
PartialColumns block is the per-level log-VMR columns times its basis.
"""
function analytic_jacobian(prof::AtmosphericProfile, linelists::Dict{GasSpecies, HITRANLinelist},
                           spec::StateVectorSpec;
                           sounder::Union{Sounder, Nothing} = nothing,
                           iasi::Sounder = IASIInstrument(), geom::ViewingGeometry = nadir_geometry(),
                           T_sfc::Union{Float64, Nothing} = nothing, ε_sfc::Float64 = 1.0,
                           observable::Symbol = :bt,
                           internal_dnu::Union{Float64, Nothing} = 0.001,
                           high_res_factor::Union{Int, Nothing} = nothing,
                           cutoff::Float64 = 25.0, lm_cutoff::Float64 = cutoff,
                           apply_continuum::Bool = true, continua = (:h2o, :co2, :co2_cia, :n2, :o2),
                           with_ils::Bool = true, apodization::Symbol = :gaussian,
                           line_mixing = nothing, T_method::Symbol = :logp_at_pcg,
                           source_function::Symbol = :cim, vmr_coupling::Bool = true,
                           backend = nothing)::Jacobian
    observable in (:bt, :radiance) || error("observable must be :bt or :radiance, got :$observable")
    length(prof.temperature) == spec.n_levels ||
        error("profile has $(length(prof.temperature)) levels, spec expects $(spec.n_levels)")
    snd = sounder === nothing ? iasi : sounder    # vmr_jacobian.jl:177: `sounder=` wins when given
    ctx = context()
    o = options(; internal_dnu, high_res_factor, cutoff, lm_cutoff, dptmn = 0.0, with_ils,
                apodization, line_mixing, T_method, source_function, apply_continuum, continua)
    keep = Any[]
    relmat_options!(o, ctx, line_mixing, linelists, keep)
    cp, _ = cprofile(ctx, prof, linelists, line_mixing, T_sfc, ε_sfc, geom.μ, keep)
    blocks = spec.vmr_blocks
    iscs = [b.basis !== nothing && size(b.basis, 2) == 1 && all(==(1.0), b.basis) for b in blocks]
    retr = Cint[Int(b.species) for b in blocks]
    cscl = Cint[c ? 1 : 0 for c in iscs]
    ss = CStateSpec(spec.include_temperature, length(blocks), pointer(retr), pointer(cscl),
                    spec.include_tsfc, spec.include_emissivity, spec.log_vmr, vmr_coupling,
                    observable === :bt ? 0 : 1)
    nl = spec.n_levels
    ndev = ccall((:irs_state_size, libirs), Cint, (Ref{CStateSpec}, Cint), ss, nl)
    grid = wavenumber_grid(snd.ν_min, snd.ν_max, snd.Δν)
    Kd = zeros(Float64, grid.n, ndev); y0 = zeros(Float64, grid.n); xd = zeros(Float64, ndev)
    cs = csounder(snd)
    GC.@preserve keep retr cscl Kd y0 xd begin
        check(ctx, ccall((:irs_jacobian, libirs), Cint,
                         (Ptr{Cvoid}, Ref{CProfile}, Ref{CSounder}, Ref{COptions}, Ref{CStateSpec},
                          Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                         ctx.h, cp, cs, o, ss, Kd, y0, xd))
    end
    K = zeros(Float64, grid.n, spec.n); x0 = zeros(Float64, spec.n)
    c = 0
    if spec.include_temperature
        K[:, spec.temp_range] .= Kd[:, 1:nl]; x0[spec.temp_range] .= xd[1:nl]; c = nl
    end
    for (b, iscol) in zip(blocks, iscs)
        if b.basis === nothing
            K[:, b.range] .= Kd[:, c+1:c+nl]; x0[b.range] .= xd[c+1:c+nl]; c += nl
        elseif iscol
            K[:, b.range] .= Kd[:, c+1:c+1]; c += 1
        else
            cols = Kd[:, c+1:c+nl]
            spec.log_vmr || (cols = cols .* Float64.(prof.vmr[b.species])')
            K[:, b.range] .= cols * b.basis; c += nl
        end
    end
    spec.include_tsfc && (K[:, spec.tsfc_index] .= Kd[:, c+1]; x0[spec.tsfc_index] = xd[c+1]; c += 1)
    spec.include_emissivity && (K[:, spec.emis_index] .= Kd[:, c+1]; x0[spec.emis_index] = xd[c+1])
    return Jacobian(K, spec, collect(grid.ν), x0, y0, observable)
end

"""
    faddeeva_voigt32(x, y; rule = :core, ctx = context()) -> Vector{Float64}

Re w(x + iy) (the Voigt H function of `faddeeva_voigt`, src/CrossSections/voigt.jl:158-160)
from the FP32 device routines of the line sums, for arrays `x`, `y >= 0`;
`rule = :core` is the CORE-range rule, `:pair` the per-pair and gradient rule.
"""
function faddeeva_voigt32(x::AbstractVector{<:Real}, y::AbstractVector{<:Real}; rule::Symbol = :core,
                          ctx::Context = context())
    length(x) == length(y) || throw(ArgumentError("x and y must have the same length"))
    rule in (:core, :pair) || throw(ArgumentError("rule must be :core or :pair"))
    xs, ys = Float64.(collect(x)), Float64.(collect(y))
    wr, wi = zeros(Float64, length(xs)), zeros(Float64, length(xs))
    check(ctx, ccall((:irs_faddeeva32, libirs), Cint,
                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Cint, Ptr{Float64}, Ptr{Float64}),
                     ctx.h, xs, ys, length(xs), rule === :core ? 0 : 1, wr, wi))
    return wr
end

"""
    Emulator

A trained and calibrated channel-transmittance emulator loaded by `irs_emu_load`
from the file the Python package writes (`ChannelODEmulator.export_binary`).  It
serves `forward_emulated`: the emulated spectra, an LBL spot check of the
calibration's channel blocks, the correction of every channel from the scene's
spot errors, and full LBL for the scenes the spot check flags.  No counterpart
in the reference.
"""
mutable struct Emulator
    mlp::Ptr{Cvoid}
    emu::Ptr{Cvoid}
    threshold::Float64
    block_c0::Vector{Int64}
    block_n::Vector{Int32}
    lists::IdDict{Any, Matrix{Int32}}     # (linelists, Y) => spot-block list handles, uploaded once
end

function load_emulator(path::AbstractString; device::Integer = parse(Int, get(ENV, "IRS_DEVICE", "0")),
                       max_batch::Integer = 4096)
    mlp = Ref{Ptr{Cvoid}}(C_NULL); emu = Ref{Ptr{Cvoid}}(C_NULL)
    thr = Ref{Float64}(0.0); nb = Ref{Cint}(0)
    c0 = zeros(Int64, 64); bn = zeros(Int32, 64)
    rc = ccall((:irs_emu_load, libirs), Cint,
               (Cstring, Cint, Cint, Ref{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}, Ref{Float64}, Cint, Ptr{Int64}, Ptr{Int32},
                Ref{Cint}),
               path, device, max_batch, mlp, emu, thr, 64, c0, bn, nb)
    rc == 0 || error("irs_emu_load($path) failed ($rc)")
    e = Emulator(mlp[], emu[], thr[], c0[1:nb[]], bn[1:nb[]], IdDict{Any, Matrix{Int32}}())
    finalizer(e) do x
        ccall((:irs_emu_destroy, libirs), Cvoid, (Ptr{Cvoid},), x.emu)
        ccall((:irs_mlp_destroy, libirs), Cvoid, (Ptr{Cvoid},), x.mlp)
    end
    return e
end

"""
    forward_emulated(prof, linelists, em::Emulator; sounder, geom, T_sfc, ε_sfc, line_mixing,
                     apply_continuum=false, threshold=em.threshold) -> (WavenumberGrid, R, BT, fell_back)

`forward_model`'s call with the learned path (`irs_forward_emulated`): the spot blocks
are the emulator's calibration blocks, each recomputed by LBL on its own sub-grid with
the lists filtered to the block ± 42 cm⁻¹; `fell_back[q]` marks the scenes recomputed
with full LBL.  `apply_continuum` and `line_mixing` as in the emulator's training
spectra.
"""
function forward_emulated(prof, linelists::Dict{GasSpecies, HITRANLinelist}, em::Emulator;
                          sounder::Sounder = IASIInstrument(), iasi::Union{Sounder, Nothing} = nothing,
                          geom = nadir_geometry(), T_sfc = nothing, ε_sfc = 1.0, line_mixing = nothing,
                          apply_continuum::Bool = false, threshold::Float64 = em.threshold,
                          cutoff::Float64 = 25.0, dptmn::Float64 = 1e-6)
    snd = iasi === nothing ? sounder : iasi
    profs = prof isa AbstractVector ? prof : [prof]
    nb = length(profs)
    pick(x, q) = x isa AbstractVector ? x[q] : x
    ctx = context()
    o = options(; internal_dnu = 0.001, high_res_factor = nothing, cutoff, lm_cutoff = cutoff, dptmn,
                with_ils = true, apodization = :gaussian, line_mixing, T_method = :logp_at_pcg,
                source_function = :cim, apply_continuum, continua = (:h2o, :co2, :co2_cia, :n2, :o2))
    keep = Any[]
    cps = Vector{CProfile}(undef, nb)
    species = GasSpecies[]
    for q in 1:nb
        cps[q], species = cprofile(ctx, profs[q], linelists, line_mixing, pick(T_sfc, q), pick(ε_sfc, q),
                                   pick(geom, q).μ, keep)
    end
    y = line_mixing isa AbstractVector ? line_mixing : nothing
    nblk = length(em.block_c0)
    # the caller's lists filtered to each block +- 42 cm^-1, uploaded on the first call with
    # these lists (and Y) and kept with the emulator: [block][species] in C order
    lists = get!(em.lists, (linelists, y, Tuple(species), snd.ν_min, snd.Δν)) do   # compared with ===
        m = Matrix{Int32}(undef, length(species), nblk)
        for b in 1:nblk
            lo = snd.ν_min + em.block_c0[b] * snd.Δν - 42.0
            hi = snd.ν_min + (em.block_c0[b] + em.block_n[b] - 1) * snd.Δν + 42.0
            for (k, sp) in enumerate(species)
                if !haskey(linelists, sp)
                    m[k, b] = -1
                    continue
                end
                ll = linelists[sp]
                sel = [lo <= l.wavenumber <= hi for l in ll.lines]
                sub = IRSounderLBL.filter_linelist(ll, lo, hi)
                yb = (sp == CO2 && y !== nothing) ? y[sel] : nothing
                m[k, b] = isempty(sub.lines) ? Int32(-1) : upload!(ctx, sub, yb)
            end
        end
        m
    end
    grid = wavenumber_grid(snd.ν_min, snd.ν_max, snd.Δν)
    R = zeros(Float64, grid.n, nb); BT = zeros(Float64, grid.n, nb)
    fb = zeros(UInt8, nb); err = zeros(Float64, nb)
    cs = csounder(snd)
    GC.@preserve keep cps lists R BT fb err begin
        rc = ccall((:irs_forward_emulated, libirs), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{CProfile}, Cint, Ref{CSounder}, Ref{COptions}, Cint, Ptr{Int64},
                    Ptr{Int32}, Ptr{Int32}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{UInt8}, Ptr{Float64}),
                   ctx.h, em.emu, cps, nb, cs, o, nblk, em.block_c0, em.block_n, lists, threshold, R, BT, fb, err)
        rc == 0 || error(unsafe_string(ccall((:irs_emu_last_error, libirs), Cstring, (Ptr{Cvoid},), em.emu)) *
                         " / " * unsafe_string(ccall((:irs_last_error, libirs), Cstring, (Ptr{Cvoid},), ctx.h)))
    end
    fell_back = fb .!= 0
    prof isa AbstractVector && return grid, R, BT, fell_back
    return grid, R[:, 1], BT[:, 1], fell_back[1]
end

end # module
