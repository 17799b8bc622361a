"""
    RippleMI355X

Julia host layer for the MI355X wave-action time step.  It adds an
architecture object, `MI355X()`, and the methods that let Ripple's
`SpectralWaveModel` / `set!` / `time_step!` / `Simulation` surface run on
libripple_mi355x.so through `ccall` (C ABI: include/ripple_mi355x.h).

Usage mirrors the reference quick start (Ripple.jl README.md:57-108):

    model = MI355XWaveModel(grid, spectral_grid; velocities=(; u, v),
                            horizontal_advection=nothing, timestepper=:RK3)
    set!(model, N = (x, y, kx, ky) -> ...)
    for _ in 1:400; time_step!(model, 0.05); end
    m0(model)

Only Libdl is needed; the grids are passed as plain vectors (centres, faces,
first spacings), so the module works with Ripple's own PolarWaveVectorGrid /
FrequencyDirectionGrid / Oceananigans RectilinearGrid objects, or with the
`SimpleGrid` helpers below when Ripple itself is not loaded.
"""
module RippleMI355X

using Libdl

export MI355X, MI355XWaveModel, set!, time_step!, interior_action, doppler_velocity,
       PseudomomentumVelocities, SpatialAveraging, update_coupling!, doppler_velocity_kappa_derivative,
       compute_pseudomomentum_cell_averages, compute_pseudomomentum_cell_integrals, set_quadruplets!,
       update_quadruplet_rates!,
       m0, root_mean_square_wavenumber, mean_direction, significant_wave_height,
       total_action, total_deep_water_energy, Simulation, run!

const LIB = Ref{Ptr{Cvoid}}(C_NULL)
const LIBPATH = joinpath(@__DIR__, "..", "ripple_amd", "libripple_mi355x.so")

struct MI355X end

# Mirrors rpl_config in include/ripple_mi355x.h.
struct RplConfig
    nx::Int32
    ny::Int32
    nkappa::Int32
    nphi::Int32
    dtype::Int32
    timestepper::Int32
    transport::Int32
    kernel::Int32
end

const RPL_F32, RPL_F64 = Int32(0), Int32(1)
const TIMESTEPPERS = Dict(:RK3 => Int32(0), :LSRK3 => Int32(1), :LowStorageRK3 => Int32(1),
                          :ForwardEuler => Int32(2))

function lib()
    if LIB[] == C_NULL
        isfile(LIBPATH) || error("$LIBPATH is missing: build it with `python ripple.jl_amd/build.py`")
        LIB[] = Libdl.dlopen(LIBPATH)
    end
    return LIB[]
end

sym(name) = Libdl.dlsym(lib(), name)

function check(code)
    code == 0 && return nothing
    msg = unsafe_string(ccall(sym(:rpl_last_error), Cstring, ()))
    code == -1 ? throw(ArgumentError(msg)) : error("ripple_mi355x error $code: $msg")
end

"""
    PseudomomentumVelocities(; q_grid=nothing)

The self-coupled current (Coupling/lagrangian_velocities.jl:65-81): the
Doppler velocity is the wave field's own pseudomomentum, recomputed from the
action before every tendency evaluation.  `q_grid` (anything with
`z_faces`) sets the vertical cells; by default the model grid's.
"""
struct PseudomomentumVelocities{G}
    q_grid::G
end
PseudomomentumVelocities(; q_grid=nothing) = PseudomomentumVelocities(q_grid)

"`SpatialAveraging(; αs=0.5, αn=0.5)` (Models/propagation_smoothing.jl:22-51), applied after every step."
struct SpatialAveraging
    αs::Float64
    αn::Float64
end
SpatialAveraging(; αs=0.5, αn=0.5) = SpatialAveraging(Float64(αs), Float64(αn))

mutable struct Clock
    time::Float64
    iteration::Int
end

mutable struct MI355XWaveModel{FT}
    ctx::Ptr{Cvoid}
    size::NTuple{4, Int}
    x::Vector{Float64}
    y::Vector{Float64}
    kappa::Vector{Float64}
    phi::Vector{Float64}
    timestepper::Symbol
    clock::Clock
end

"""
    MI355XWaveModel(grid, spectral_grid; velocities=nothing, horizontal_advection=:WENO,
                    spectral_advection=:WENO, depth=nothing, timestepper=:ForwardEuler,
                    float_type=Float64, gravity=9.81)

`grid` needs `x_faces`, `y_faces`, `z_faces` vectors (Oceananigans: `xfaces(grid)` ...);
`spectral_grid` needs `κ`, `κ_faces`, `φ`, `φ_faces`.  Dispatch follows
Models/tendencies.jl:18-109.  `depth === nothing` is InfiniteDepth().  `topology`
is the grid's (x, y) topology, `:Periodic` or `:Bounded` (Oceananigans: `topology(grid)`).
"""
function MI355XWaveModel(grid, spectral_grid; velocities=nothing, horizontal_advection=:WENO,
                         spectral_advection=:WENO, depth=nothing, timestepper=:ForwardEuler,
                         float_type=Float64, gravity=9.81, kernel=0, propagation_smoothing=nothing,
                         topology=(:Periodic, :Periodic))
    xf, yf, zf = collect(Float64, grid.x_faces), collect(Float64, grid.y_faces), collect(Float64, grid.z_faces)
    κ, κf = collect(Float64, spectral_grid.κ), collect(Float64, spectral_grid.κ_faces)
    φ, φf = collect(Float64, spectral_grid.φ), collect(Float64, spectral_grid.φ_faces)
    Nx, Ny, Nκ, Nφ = length(xf) - 1, length(yf) - 1, length(κ), length(φ)
    haskey(TIMESTEPPERS, timestepper) || throw(ArgumentError("unsupported timestepper $timestepper"))
    transport = if velocities !== nothing && spectral_advection !== nothing
        Int32(2)
    elseif horizontal_advection === nothing
        Int32(0)
    elseif Nκ * Nφ > 1
        Int32(1)
    else
        throw(ArgumentError("single-bin transport stays on the Oceananigans per-bin path"))
    end
    cfg = RplConfig(Nx, Ny, Nκ, Nφ, float_type === Float64 ? RPL_F64 : RPL_F32,
                    TIMESTEPPERS[timestepper], transport, Int32(kernel))
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall(sym(:rpl_create), Cint, (Ref{Ptr{Cvoid}}, Ref{RplConfig}), ctx, cfg))
    model = MI355XWaveModel{float_type}(ctx[], (Nx, Ny, Nκ, Nφ), xf, yf, κ, φ, timestepper, Clock(0.0, 0))
    finalizer(m -> ccall(sym(:rpl_destroy), Cint, (Ptr{Cvoid},), m.ctx), model)
    depth_kind, d = depth === nothing ? (Int32(0), 0.0) : (Int32(1), Float64(depth))
    check(ccall(sym(:rpl_set_geometry), Cint,
                (Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble},
                 Int32, Cdouble, Cdouble),
                model.ctx, xf[2] - xf[1], yf[2] - yf[1], κ, κf, φ, φf, depth_kind, d, gravity))
    if velocities isa PseudomomentumVelocities
        qz = velocities.q_grid === nothing ? zf : collect(Float64, velocities.q_grid.z_faces)
        length(qz) >= 2 || throw(ArgumentError("PseudomomentumVelocities needs vertical cells (a q_grid or a 3-D grid)"))
        q_depth = depth === nothing ? qz[end] - qz[1] : Float64(depth)
        check(ccall(sym(:rpl_set_pseudomomentum_coupling), Cint, (Ptr{Cvoid}, Int32, Ptr{Cdouble}, Cdouble),
                    model.ctx, length(qz) - 1, qz, q_depth))
    elseif velocities !== nothing
        u = collect(Float64, velocities.u)
        v = collect(Float64, velocities.v)
        Nz = length(zf) - 1
        size(u) == (Nx, Ny, Nz) || throw(ArgumentError("velocity arrays must have size ($Nx, $Ny, $Nz)"))
        q_depth = depth === nothing ? zf[end] - zf[1] : Float64(depth)
        check(ccall(sym(:rpl_set_current), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble),
                    model.ctx, Nz, zf, u, v, q_depth))
    end
    check(ccall(sym(:rpl_set_topology), Cint, (Ptr{Cvoid}, Int32, Int32),
                model.ctx, Int32(topology[1] === :Periodic), Int32(topology[2] === :Periodic)))
    if propagation_smoothing isa SpatialAveraging
        check(ccall(sym(:rpl_set_propagation_smoothing), Cint, (Ptr{Cvoid}, Int32, Cdouble, Cdouble),
                    model.ctx, Int32(1), propagation_smoothing.αs, propagation_smoothing.αn))
    end
    return model
end

xnodes(m) = (m.x[1:end-1] .+ m.x[2:end]) ./ 2
ynodes(m) = (m.y[1:end-1] .+ m.y[2:end]) ./ 2

"`set!(model; N)`: N is a number, an (Nx, Ny, Nκ, Nφ) array, or `f(x, y, kx, ky)` (ProductFields/set.jl)."
function set!(model::MI355XWaveModel{FT}; N=nothing) where FT
    N === nothing && return model
    vals = if N isa Number
        fill(FT(N), model.size)
    elseif N isa AbstractArray
        size(N) == model.size || throw(ArgumentError("array size $(size(N)) does not match field size $(model.size)"))
        convert(Array{FT, 4}, N)
    else
        xs, ys = xnodes(model), ynodes(model)
        [FT(N(xs[i], ys[j], model.kappa[m] * cos(model.phi[n]), model.kappa[m] * sin(model.phi[n])))
         for i in eachindex(xs), j in eachindex(ys), m in eachindex(model.kappa), n in eachindex(model.phi)]
    end
    check(ccall(sym(:rpl_set_action), Cint, (Ptr{Cvoid}, Ptr{FT}), model.ctx, vals))
    return model
end

"`time_step!(model, dt)` (Models/time_step.jl:119-170), run on the device."
function time_step!(model::MI355XWaveModel, dt; nsteps::Integer=1)
    dt > 0 || throw(ArgumentError("time step must be positive"))
    check(ccall(sym(:rpl_time_step), Cint, (Ptr{Cvoid}, Cdouble, Int32), model.ctx, dt, nsteps))
    t, it = Ref{Cdouble}(0), Ref{Int64}(0)
    check(ccall(sym(:rpl_get_clock), Cint, (Ptr{Cvoid}, Ref{Cdouble}, Ref{Int64}), model.ctx, t, it))
    model.clock.time, model.clock.iteration = t[], it[]
    return model
end

function interior_action(model::MI355XWaveModel{FT}) where FT
    N = Array{FT}(undef, model.size)
    check(ccall(sym(:rpl_get_action), Cint, (Ptr{Cvoid}, Ptr{FT}), model.ctx, N))
    return N
end

function doppler_velocity(model::MI355XWaveModel)
    Nx, Ny, Nκ, _ = model.size
    Ux, Uy = zeros(Nx, Ny, Nκ), zeros(Nx, Ny, Nκ)
    check(ccall(sym(:rpl_get_doppler), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}), model.ctx, Ux, Uy))
    return Ux, Uy
end

"`update_coupling!(model)` (Coupling/current_coupling.jl:123-148)."
update_coupling!(model::MI355XWaveModel) = (check(ccall(sym(:rpl_update_coupling), Cint, (Ptr{Cvoid},), model.ctx)); model)

function doppler_velocity_kappa_derivative(model::MI355XWaveModel)
    Nx, Ny, Nκ, _ = model.size
    dUx, dUy = zeros(Nx, Ny, Nκ), zeros(Nx, Ny, Nκ)
    check(ccall(sym(:rpl_get_doppler_kappa_derivative), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}),
                model.ctx, dUx, dUy))
    return dUx, dUy
end

"Pseudomomentum cell averages (or integrals) on the vertical cells `z_faces` (Coupling/pseudomomentum.jl:162-210)."
function _pseudomomentum(model::MI355XWaveModel, z_faces, depth, average)
    zf = collect(Float64, z_faces)
    Nx, Ny = model.size[1], model.size[2]
    px, py = zeros(Nx, Ny, length(zf) - 1), zeros(Nx, Ny, length(zf) - 1)
    check(ccall(sym(:rpl_pseudomomentum), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Cdouble}, Cdouble, Int32, Ptr{Cdouble}, Ptr{Cdouble}),
                model.ctx, length(zf) - 1, zf, Float64(depth), Int32(average), px, py))
    return px, py
end
compute_pseudomomentum_cell_averages(model, z_faces, depth) = _pseudomomentum(model, z_faces, depth, true)
compute_pseudomomentum_cell_integrals(model, z_faces, depth) = _pseudomomentum(model, z_faces, depth, false)

"""
    set_quadruplets!(model, interactions, frequency; power=1.0, resonance_tolerance=1e-12)

Installs `DiscreteInteractionApproximation(interactions; power,
resonance_tolerance)` with the `ListedQuadruplets` method
(Sources/Sources.jl:378-428, 1268-1385).  `interactions` are objects with
Ripple's `QuadrupletTransferInteraction` fields (1-based `donor1_m`, …,
`receiver2_n`, `rate`); `frequency` is the grid's frequency centres.  A
`rate` may be a number, an `(Nx, Ny)` array or a function of `(x, y, t)`,
`(x, y)` or `(t)` (Ripple's `source_value`, Sources.jl:1450-1467); the
non-constant ones go to the device as a per-column field evaluated at
`model.clock.time` (`update_quadruplet_rates!` re-evaluates it, before each
step for a rate that depends on time).  The library runs the reference's
index, resonance and rate checks and throws on the first failure.
"""
function set_quadruplets!(model::MI355XWaveModel, interactions, frequency; power=1.0, resonance_tolerance=1e-12)
    bins = Int32[]
    rates = Float64[]
    for q in interactions
        append!(bins, Int32.((q.donor1_m, q.donor1_n, q.donor2_m, q.donor2_n,
                              q.receiver1_m, q.receiver1_n, q.receiver2_m, q.receiver2_n)) .- Int32(1))
        push!(rates, q.rate isa Number ? Float64(q.rate) : 0.0)
    end
    check(ccall(sym(:rpl_set_quadruplets), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Cdouble}, Ptr{Cdouble}, Cdouble, Cdouble),
                model.ctx, length(rates), bins, rates, collect(Float64, frequency), Float64(power),
                Float64(resonance_tolerance)))
    all(q -> q.rate isa Number, interactions) || update_quadruplet_rates!(model, interactions)
    return model
end

function _source_value(value, x, y, t)
    value isa Number && return Float64(value)
    applicable(value, x, y, t) && return Float64(value(x, y, t))
    applicable(value, x, y) && return Float64(value(x, y))
    applicable(value, t) && return Float64(value(t))
    throw(ArgumentError("callable source parameters must accept (x, y, t), (x, y), or (t)"))
end

"""
    update_quadruplet_rates!(model, interactions)

Evaluates every interaction's rate at the cell centres and `model.clock.time`
and uploads them as the per-column rate field (`rpl_set_quadruplet_rate_field`).
"""
function update_quadruplet_rates!(model::MI355XWaveModel, interactions)
    Nx, Ny = model.size[1], model.size[2]
    xc, yc = xnodes(model), ynodes(model)
    t = model.clock.time
    field = zeros(Nx, Ny, length(interactions))
    for (k, q) in enumerate(interactions), j in 1:Ny, i in 1:Nx
        field[i, j, k] = q.rate isa AbstractArray ? Float64(q.rate[i, j]) : _source_value(q.rate, xc[i], yc[j], t)
    end
    check(ccall(sym(:rpl_set_quadruplet_rate_field), Cint, (Ptr{Cvoid}, Ptr{Cdouble}), model.ctx, field))
    return model
end

function _columns(model, which)
    Nx, Ny = model.size[1], model.size[2]
    out = [zeros(Nx, Ny) for _ in 1:3]
    ptrs = [which[k] ? pointer(out[k]) : Ptr{Cdouble}(C_NULL) for k in 1:3]
    GC.@preserve out check(ccall(sym(:rpl_diagnostics), Cint, (Ptr{Cvoid}, Ptr{Cdouble}, Ptr{Cdouble}, Ptr{Cdouble}),
                                 model.ctx, ptrs...))
    return out
end

m0(model::MI355XWaveModel) = _columns(model, (true, false, false))[1]
root_mean_square_wavenumber(model::MI355XWaveModel) = _columns(model, (false, true, false))[2]
mean_direction(model::MI355XWaveModel) = _columns(model, (false, false, true))[3]
significant_wave_height(model::MI355XWaveModel) = 4 .* sqrt.(m0(model))

function _totals(model)
    a, e = Ref{Cdouble}(0), Ref{Cdouble}(0)
    check(ccall(sym(:rpl_conservation), Cint, (Ptr{Cvoid}, Ref{Cdouble}, Ref{Cdouble}), model.ctx, a, e))
    return a[], e[]
end

total_action(model::MI355XWaveModel) = _totals(model)[1]
total_deep_water_energy(model::MI355XWaveModel) = _totals(model)[2]

struct Simulation{M}
    model::M
    Δt::Float64
    stop_iteration::Int
end

Simulation(model::MI355XWaveModel; Δt, stop_iteration) = Simulation(model, Float64(Δt), Int(stop_iteration))

function run!(sim::Simulation)
    n = sim.stop_iteration - sim.model.clock.iteration
    n > 0 && time_step!(sim.model, sim.Δt; nsteps=n)
    check(ccall(sym(:rpl_synchronize), Cint, (Ptr{Cvoid},), sim.model.ctx))
    return sim
end

end # module
