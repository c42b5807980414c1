# MathOptInterface front-end over libpdcs_hip.so.
#
# Same model class, options and result attributes as the reference's
# `PDCS_GPU.Optimizer` (src/pdcs_gpu/MOI_wrapper/MOI_wrapper.jl:165-873):
#   min/max c'x + c0  s.t.  A x + b in K,  l <= x <= u,
# K a product of Zeros, Nonnegatives, SecondOrderCone, RotatedSecondOrderCone,
# ExponentialCone and DualExponentialCone blocks (VectorAffineFunction rows),
# bounds from VariableIndex constraints.  Rows are grouped by cone type in that
# order (the order the C ABI's cone descriptor expects); the library then
# solves  min c'x  s.t.  G x - h in K  with G = A, h = -b (a maximization is
# solved as the minimization of -c'x).
#
# Deliberate differences from the reference wrapper:
#   * exit codes 2/3 (primal infeasible) map to INFEASIBLE and 4/5 (dual
#     infeasible) to DUAL_INFEASIBLE (the reference maps 3 to DUAL_INFEASIBLE
#     and 4 to INFEASIBLE, :760-770);
#   * primal / dual status: a certificate only for the side that has one, and
#     NO_SOLUTION for 1/6/7 (the reference compares the code with a tuple,
#     `== (1, 7)`, which is never true, :828);
#   * ConstraintPrimal returns the primal slack A x + b of the rows (the
#     reference indexes an n-long vector with row indices, :840);
#   * VariableIndex bound duals come from the reduced cost c - G'y.

import MathOptInterface as MOI

MOI.Utilities.@product_of_sets(
    _ConeRows,
    MOI.Zeros,
    MOI.Nonnegatives,
    MOI.SecondOrderCone,
    MOI.RotatedSecondOrderCone,
    MOI.ExponentialCone,
    MOI.DualExponentialCone,
)

const _ModelCache = MOI.Utilities.GenericModel{
    Float64,
    MOI.Utilities.ObjectiveContainer{Float64},
    MOI.Utilities.VariablesContainer{Float64},
    MOI.Utilities.MatrixOfConstraints{
        Float64,
        MOI.Utilities.MutableSparseMatrixCSC{Float64,Int,MOI.Utilities.OneBasedIndexing},
        Vector{Float64},
        _ConeRows{Float64},
    },
}

const _ConeSets = Union{MOI.Zeros,MOI.Nonnegatives,MOI.SecondOrderCone,MOI.RotatedSecondOrderCone,
                        MOI.ExponentialCone,MOI.DualExponentialCone}
const _BoundSets = Union{MOI.EqualTo{Float64},MOI.LessThan{Float64},MOI.GreaterThan{Float64},
                         MOI.Interval{Float64}}

Base.@kwdef mutable struct MOISolution
    primal::Vector{Float64} = Float64[]
    dual::Vector{Float64} = Float64[]
    slack::Vector{Float64} = Float64[]        # A x + b, one entry per row
    reduced::Vector{Float64} = Float64[]      # c - G'y of the minimization, one per variable
    exit_code::Int = -1                       # -1: no solve yet
    exit_status::Symbol = :unknown
    objective_value::Float64 = NaN
    dual_objective_value::Float64 = NaN
    solve_time_sec::Float64 = NaN
    iterations::Int = 0
    objective_constant::Float64 = 0.0
end

"""
    Optimizer()

MathOptInterface optimizer solving on the GPU through libpdcs_hip.so.
Raw options (MOI.RawOptimizerAttribute): `rel_tol`, `abs_tol` (default 1e-6),
`time_limit_secs` (1000), `precision` (`:fp64` or `:mixed`), `verbose`,
`max_outer_iter`, `max_inner_iter`, `check_terminate_freq`,
`use_kkt_restart`, `kkt_restart_freq`, `use_duality_gap_restart`,
`duality_gap_restart_freq`, `device`.
"""
mutable struct Optimizer <: MOI.AbstractOptimizer
    rows::Union{Nothing,_ConeRows{Float64}}
    sol::MOISolution
    silent::Bool
    options::Dict{Symbol,Any}
    Optimizer() = new(nothing, MOISolution(), false, Dict{Symbol,Any}())
end

MOI.get(::Optimizer, ::MOI.SolverName) = "PDCS_GPU_AMD"
MOI.is_empty(o::Optimizer) = o.rows === nothing
function MOI.empty!(o::Optimizer)
    o.rows = nothing
    o.sol = MOISolution()
    return
end

MOI.supports(::Optimizer, ::MOI.RawOptimizerAttribute) = true
MOI.set(o::Optimizer, p::MOI.RawOptimizerAttribute, v) = (o.options[Symbol(p.name)] = v; nothing)
MOI.get(o::Optimizer, p::MOI.RawOptimizerAttribute) = o.options[Symbol(p.name)]

MOI.supports(::Optimizer, ::MOI.Silent) = true
MOI.set(o::Optimizer, ::MOI.Silent, v::Bool) = (o.silent = v; nothing)
MOI.get(o::Optimizer, ::MOI.Silent) = o.silent

MOI.supports(::Optimizer, ::MOI.TimeLimitSec) = true
MOI.set(o::Optimizer, ::MOI.TimeLimitSec, t::Real) = (o.options[:time_limit_secs] = Float64(t); nothing)
MOI.set(o::Optimizer, ::MOI.TimeLimitSec, ::Nothing) = (delete!(o.options, :time_limit_secs); nothing)
MOI.get(o::Optimizer, ::MOI.TimeLimitSec) = get(o.options, :time_limit_secs, nothing)

MOI.supports(::Optimizer, ::Union{MOI.ObjectiveSense,MOI.ObjectiveFunction{MOI.ScalarAffineFunction{Float64}}}) = true
MOI.supports(::Optimizer, ::MOI.VariablePrimalStart, ::Type{MOI.VariableIndex}) = true
MOI.supports(::Optimizer, ::MOI.ConstraintDualStart, ::Type{<:MOI.ConstraintIndex}) = true
MOI.supports_constraint(::Optimizer, ::Type{MOI.VectorAffineFunction{Float64}}, ::Type{<:_ConeSets}) = true
MOI.supports_constraint(::Optimizer, ::Type{MOI.VariableIndex}, ::Type{<:_BoundSets}) = true

_dims(Ab, ::Type{S}) where {S} =
    Int[MOI.dimension(MOI.get(Ab, MOI.ConstraintSet(), ci))
        for ci in MOI.get(Ab, MOI.ListOfConstraintIndices{MOI.VectorAffineFunction{Float64},S}())]

function MOI.optimize!(dest::Optimizer, src::MOI.ModelLike)
    cache = MOI.Utilities.UniversalFallback(_ModelCache())
    index_map = MOI.copy_to(cache, src)
    _solve!(dest, cache)
    return index_map, false
end

function _solve!(dest::Optimizer, src::MOI.Utilities.UniversalFallback{_ModelCache})
    MOI.empty!(dest)
    for (F, S) in keys(src.constraints)      # anything the typed cache did not take
        throw(MOI.UnsupportedConstraint{F,S}())
    end
    Ab = src.model.constraints
    A = Ab.coefficients
    n, m = A.n, A.m
    m > 0 || error("The constraint matrix G is empty, PDHG solver cannot be applied.")
    # objective
    max_sense = MOI.get(src, MOI.ObjectiveSense()) == MOI.MAX_SENSE
    c = zeros(n)
    c0 = 0.0
    for attr in MOI.get(src, MOI.ListOfModelAttributesSet())
        if attr == MOI.ObjectiveSense() || attr == MOI.Name()
            continue
        elseif attr == MOI.ObjectiveFunction{MOI.ScalarAffineFunction{Float64}}()
            f = MOI.get(src, attr)
            c0 = MOI.constant(f)
            for t in f.terms
                c[t.variable.value] += (max_sense ? -1.0 : 1.0) * t.coefficient
            end
        else
            throw(MOI.UnsupportedAttribute(attr))
        end
    end
    # warm start (primal values and row duals), zero where not given
    x0 = zeros(n); y0 = zeros(m); warm = false
    for attr in MOI.get(src, MOI.ListOfVariableAttributesSet())
        if attr == MOI.VariablePrimalStart()
            warm = true
            for (i, vi) in enumerate(MOI.get(src, MOI.ListOfVariableIndices()))
                x0[i] = something(MOI.get(src, attr, vi), 0.0)
            end
        elseif attr != MOI.VariableName()
            throw(MOI.UnsupportedAttribute(attr))
        end
    end
    for (F, S) in MOI.get(src, MOI.ListOfConstraintTypesPresent())
        for attr in MOI.get(src, MOI.ListOfConstraintAttributesSet{F,S}())
            if attr == MOI.ConstraintDualStart() && F == MOI.VectorAffineFunction{Float64}
                warm = true
                for ci in MOI.get(src, MOI.ListOfConstraintIndices{F,S}())
                    v = MOI.get(src, attr, ci)
                    v === nothing || (y0[MOI.Utilities.rows(Ab, ci)] .= v)
                end
            elseif attr != MOI.ConstraintName() && attr != MOI.ConstraintPrimalStart()
                throw(MOI.UnsupportedAttribute(attr))
            end
        end
    end
    G = SparseMatrixCSC{Float64,Int}(m, n, collect(Int, A.colptr), collect(Int, A.rowval), copy(A.nzval))
    h = -Ab.constants
    bl = copy(src.model.variables.lower)
    bu = copy(src.model.variables.upper)
    opt = dest.options
    verbose = dest.silent ? 0 : get(opt, :verbose, 0)
    res = rpdhg_gpu_solve(
        n = n, m = m, nb = n, c = c, G = G, h = h,
        mGzero = sum(_dims(Ab, MOI.Zeros); init = 0),
        mGnonnegative = sum(_dims(Ab, MOI.Nonnegatives); init = 0),
        socG = _dims(Ab, MOI.SecondOrderCone),
        rsocG = _dims(Ab, MOI.RotatedSecondOrderCone),
        expG = length(_dims(Ab, MOI.ExponentialCone)),
        dual_expG = length(_dims(Ab, MOI.DualExponentialCone)),
        bl = bl, bu = bu, soc_x = Int[], rsoc_x = Int[],
        primal_sol = x0, dual_sol = y0, warm_start = warm,
        rel_tol = Float64(get(opt, :rel_tol, 1e-6)),
        abs_tol = Float64(get(opt, :abs_tol, 1e-6)),
        time_limit = Float64(get(opt, :time_limit_secs, 1000.0)),
        max_outer_iter = get(opt, :max_outer_iter, 10000),
        max_inner_iter = get(opt, :max_inner_iter, 500000),
        check_terminate_freq = get(opt, :check_terminate_freq, 2000),
        use_kkt_restart = get(opt, :use_kkt_restart, false),
        kkt_restart_freq = get(opt, :kkt_restart_freq, 2000),
        use_duality_gap_restart = get(opt, :use_duality_gap_restart, true),
        duality_gap_restart_freq = get(opt, :duality_gap_restart_freq, 2000),
        precision = Symbol(get(opt, :precision, :fp64)),
        verbose = verbose,
        device = get(opt, :device, 0),
    )
    x = res.x.recovered_primal.primal_sol.x
    y = res.y.recovered_dual.dual_sol.y
    sgn = max_sense ? -1.0 : 1.0
    dest.rows = deepcopy(Ab.sets)
    dest.sol = MOISolution(
        primal = x, dual = y, slack = G * x - h, reduced = res.y.slack.primal_sol.x,
        exit_code = res.info.exit_code, exit_status = res.info.exit_status,
        objective_value = sgn * res.info.pObj, dual_objective_value = sgn * res.info.dObj,
        solve_time_sec = res.info.time, iterations = res.info.iter, objective_constant = c0,
    )
    return
end

"""
    PDHGIterations()

The number of PDHG iterations of the last solve.
"""
struct PDHGIterations <: MOI.AbstractModelAttribute end
MOI.is_set_by_optimize(::PDHGIterations) = true
MOI.get(o::Optimizer, ::PDHGIterations) = o.sol.iterations

MOI.get(o::Optimizer, ::MOI.SolveTimeSec) = o.sol.solve_time_sec
MOI.get(o::Optimizer, ::MOI.RawStatusString) = String(o.sol.exit_status)
MOI.get(o::Optimizer, ::MOI.ResultCount) = o.sol.exit_code < 0 ? 0 : 1

function MOI.get(o::Optimizer, ::MOI.TerminationStatus)
    s = o.sol.exit_code
    s < 0 && return MOI.OPTIMIZE_NOT_CALLED
    s == 0 && return MOI.OPTIMAL
    s in (1, 7) && return MOI.ITERATION_LIMIT
    s in (2, 3) && return MOI.INFEASIBLE
    s in (4, 5) && return MOI.DUAL_INFEASIBLE
    s == 6 && return MOI.TIME_LIMIT
    return MOI.NUMERICAL_ERROR                 # 8 (NaN), 9 (Inf)
end

function MOI.get(o::Optimizer, attr::MOI.PrimalStatus)
    attr.result_index > MOI.get(o, MOI.ResultCount()) && return MOI.NO_SOLUTION
    s = o.sol.exit_code
    s == 0 && return MOI.FEASIBLE_POINT
    s in (4, 5) && return MOI.INFEASIBILITY_CERTIFICATE   # primal ray
    s in (2, 3, 8, 9) && return MOI.NO_SOLUTION
    return MOI.UNKNOWN_RESULT_STATUS                       # 1, 6, 7: last iterate
end

function MOI.get(o::Optimizer, attr::MOI.DualStatus)
    attr.result_index > MOI.get(o, MOI.ResultCount()) && return MOI.NO_SOLUTION
    s = o.sol.exit_code
    s == 0 && return MOI.FEASIBLE_POINT
    s in (2, 3) && return MOI.INFEASIBILITY_CERTIFICATE   # dual ray
    s in (4, 5, 8, 9) && return MOI.NO_SOLUTION
    return MOI.UNKNOWN_RESULT_STATUS
end

function MOI.get(o::Optimizer, attr::MOI.ObjectiveValue)
    MOI.check_result_index_bounds(o, attr)
    v = o.sol.objective_value
    MOI.Utilities.is_ray(MOI.get(o, MOI.PrimalStatus())) || (v += o.sol.objective_constant)
    return v
end

function MOI.get(o::Optimizer, attr::MOI.DualObjectiveValue)
    MOI.check_result_index_bounds(o, attr)
    v = o.sol.dual_objective_value
    MOI.Utilities.is_ray(MOI.get(o, MOI.DualStatus())) || (v += o.sol.objective_constant)
    return v
end

function MOI.get(o::Optimizer, attr::MOI.VariablePrimal, vi::MOI.VariableIndex)
    MOI.check_result_index_bounds(o, attr)
    return o.sol.primal[vi.value]
end

function MOI.get(o::Optimizer, attr::MOI.ConstraintPrimal,
                 ci::MOI.ConstraintIndex{MOI.VectorAffineFunction{Float64},S}) where {S}
    MOI.check_result_index_bounds(o, attr)
    return o.sol.slack[MOI.Utilities.rows(o.rows, ci)]
end

function MOI.get(o::Optimizer, attr::MOI.ConstraintPrimal, ci::MOI.ConstraintIndex{MOI.VariableIndex,S}) where {S}
    MOI.check_result_index_bounds(o, attr)
    return o.sol.primal[ci.value]
end

function MOI.get(o::Optimizer, attr::MOI.ConstraintDual,
                 ci::MOI.ConstraintIndex{MOI.VectorAffineFunction{Float64},S}) where {S}
    MOI.check_result_index_bounds(o, attr)
    return o.sol.dual[MOI.Utilities.rows(o.rows, ci)]
end

# bound duals: the part of the reduced cost c - G'y on the side of the bound
function MOI.get(o::Optimizer, attr::MOI.ConstraintDual, ci::MOI.ConstraintIndex{MOI.VariableIndex,S}) where {S}
    MOI.check_result_index_bounds(o, attr)
    r = o.sol.reduced[ci.value]
    S <: MOI.GreaterThan && return max(r, 0.0)
    S <: MOI.LessThan && return min(r, 0.0)
    return r
end
