"""
    PDCS_GPU_AMD

Julia host shim over libpdcs_hip.so (include/pdcs.h).  It keeps the solve entry
points and keyword arguments of the reference's `PDCS_GPU` module
(`rpdhg_gpu_solve`, src/pdcs_gpu/rpdhg_alg_gpu_gen.jl:1219; `rpdhg_gpu_solve_input_gpu_data`,
:387) and returns a `Solution`-shaped value exposing the fields the reference's
MOI wrapper reads (src/pdcs_gpu/MOI_wrapper/MOI_wrapper.jl:648-659, 700-710):
`sol.x.recovered_primal.primal_sol.x`, `sol.y.recovered_dual.dual_sol.y`,
`sol.y.slack.primal_sol.x`, `sol.info.{exit_code, exit_status, pObj, dObj, time, iter}`.

The MOI wrapper's `SparseMatrixCSC` (1-based CSC) is converted to the 0-based
CSR the C ABI takes.  This file is not exercised by the repository's tests
(there is no Julia toolchain in the build image).
"""
module PDCS_GPU_AMD

using SparseArrays

const libpdcs = Ref{String}(normpath(joinpath(@__DIR__, "..", "pdcs_amd", "libpdcs_hip.so")))

struct PdcsCones
    mzero::Int64
    mnonneg::Int64
    nsoc::Int64
    soc::Ptr{Int64}
    nrsoc::Int64
    rsoc::Ptr{Int64}
    nexp::Int64
    ndualexp::Int64
end

mutable struct PdcsParams
    max_outer_iter::Int64
    max_inner_iter::Int64
    abs_tol::Float64
    rel_tol::Float64
    time_limit::Float64
    check_terminate_freq::Int64
    duality_gap_restart_freq::Int64
    kkt_restart_freq::Int64
    use_kkt_restart::Int32
    use_duality_gap_restart::Int32
    use_preconditioner::Int32
    verbose::Int32
    warm_start::Int32
    precision::Int32
    polish::Int32
    bisect_intent::Int32
    code_primal_denominator::Int32
    use_graph::Int32
    mixed_check_freq::Int64
    polish_check_freq::Int64
    mixed_switch_tol::Float64
    autotune::Int32
    reserved::Int32
    PdcsParams() = new()
end

mutable struct PdcsResult
    exit_code::Int32
    sequence::Int32
    iter::Int64
    iter_fp32::Int64
    restarts::Int64
    pobj::Float64
    dobj::Float64
    rel_primal_res::Float64
    rel_dual_res::Float64
    rel_gap::Float64
    abs_primal_res::Float64
    abs_dual_res::Float64
    abs_gap::Float64
    l2_rel_primal_res::Float64
    l2_rel_dual_res::Float64
    rel_primal_res_code::Float64
    omega::Float64
    time_setup::Float64
    time_solve::Float64
    time_fp32::Float64
    time_kkt::Float64
    time_restart::Float64
    sched::NTuple{4, Int32}
    PdcsResult() = new()
end

const EXIT_STATUS = Dict(0 => :optimal, 1 => :max_iter, 2 => :primal_infeasible_low_acc,
    3 => :primal_infeasible_high_acc, 4 => :dual_infeasible_low_acc, 5 => :dual_infeasible_high_acc,
    6 => :time_limit, 7 => :continue, 8 => :numerical_error_nan, 9 => :numerical_error_inf)

function check(rc::Cint, ctx::Ptr{Cvoid})
    if rc != 0
        msg = unsafe_string(ccall((:pdcs_last_error, libpdcs[]), Cstring, (Ptr{Cvoid},), ctx))
        error("PDCS_GPU_AMD: $msg (code $rc)")
    end
end

"""CSC (1-based, Julia) -> CSR (0-based, Int32) of the same matrix."""
function csr0(G::SparseMatrixCSC{Float64})
    T = SparseMatrixCSC(transpose(G))          # CSC of G' is the CSR of G
    return Int32.(T.colptr .- 1), Int32.(T.rowval .- 1), Vector{Float64}(T.nzval)
end

function rpdhg_gpu_solve(;
    n::Integer, m::Integer, nb::Integer, c::Vector{Float64}, G::SparseMatrixCSC{Float64}, h::Vector{Float64},
    mGzero::Integer, mGnonnegative::Integer, socG::Vector{<:Integer}, rsocG::Vector{<:Integer},
    expG::Integer, dual_expG::Integer, bl::Vector{Float64}, bu::Vector{Float64},
    soc_x::Vector{<:Integer}, rsoc_x::Vector{<:Integer}, exp_x::Integer = 0, dual_exp_x::Integer = 0,
    rescaling_method::Symbol = :ruiz_pock_chambolle, use_preconditioner::Bool = true,
    use_adaptive_restart::Bool = true, use_adaptive_step_size_weight::Bool = true,
    use_aggressive::Bool = true, use_accelerated::Bool = false, use_resolving::Bool = true,
    primal_sol::Vector{Float64} = zeros(n), dual_sol::Vector{Float64} = zeros(m), warm_start::Bool = false,
    max_outer_iter::Integer = 10000, max_inner_iter::Integer = 500000, abs_tol::Float64 = 1e-6,
    rel_tol::Float64 = 1e-6, eps_primal_infeasible_low_acc::Float64 = 1e-12,
    eps_dual_infeasible_low_acc::Float64 = 1e-12, eps_primal_infeasible_high_acc::Float64 = 1e-16,
    eps_dual_infeasible_high_acc::Float64 = 1e-16, print_freq::Integer = 2000,
    check_terminate_freq::Integer = 2000, verbose::Integer = 1, time_limit::Float64 = Inf,
    method::Symbol = :average, logfile_name = nothing, use_kkt_restart::Bool = false,
    kkt_restart_freq::Integer = 2000, use_duality_gap_restart::Bool = true,
    duality_gap_restart_freq::Integer = 2000, precision::Symbol = :fp64, device::Integer = 0)

    (isempty(soc_x) && isempty(rsoc_x) && exp_x == 0 && dual_exp_x == 0) ||
        throw(ArgumentError("primal-side cones are not supported"))
    (rescaling_method == :ruiz_pock_chambolle && use_preconditioner) ||
        throw(ArgumentError("only :ruiz_pock_chambolle with the preconditioner is supported"))
    (use_adaptive_restart && use_adaptive_step_size_weight && use_resolving && use_aggressive &&
        !use_accelerated && method == :average) ||
        throw(ArgumentError("only the aggressive restarted adaptive main loop is supported"))

    rowptr, colidx, vals = csr0(G)
    soc = Vector{Int64}(socG)
    rsoc = Vector{Int64}(rsocG)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:pdcs_ctx_create, libpdcs[]), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, ctx), C_NULL)
    prob = Ref{Ptr{Cvoid}}(C_NULL)
    res = PdcsResult()
    x = zeros(n); y = zeros(m); s = zeros(n)
    try
        GC.@preserve soc rsoc begin
            cones = Ref(PdcsCones(mGzero, mGnonnegative, length(soc), pointer(soc), length(rsoc), pointer(rsoc),
                                  expG, dual_expG))
            check(ccall((:pdcs_problem_create_host, libpdcs[]), Cint,
                (Ptr{Cvoid}, Int64, Int64, Int64, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64},
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{PdcsCones}, Ptr{Ptr{Cvoid}}),
                ctx[], n, m, nb, c, rowptr, colidx, vals, h, bl, bu, cones, prob), ctx[])
        end
        prm = PdcsParams()
        ccall((:pdcs_params_default, libpdcs[]), Cvoid, (Ref{PdcsParams},), prm)
        prm.max_outer_iter = max_outer_iter; prm.max_inner_iter = max_inner_iter
        prm.abs_tol = abs_tol; prm.rel_tol = rel_tol; prm.time_limit = time_limit
        prm.check_terminate_freq = check_terminate_freq
        prm.use_kkt_restart = use_kkt_restart; prm.kkt_restart_freq = kkt_restart_freq
        prm.use_duality_gap_restart = use_duality_gap_restart
        prm.duality_gap_restart_freq = duality_gap_restart_freq
        prm.verbose = verbose > 1 ? 1 : 0
        prm.precision = precision == :mixed ? 1 : 0
        prm.warm_start = warm_start ? 1 : 0
        check(ccall((:pdcs_solve, libpdcs[]), Cint,
            (Ptr{Cvoid}, Ref{PdcsParams}, Ptr{Float64}, Ptr{Float64}, Ref{PdcsResult}),
            prob[], prm, warm_start ? primal_sol : C_NULL, warm_start ? dual_sol : C_NULL, res), ctx[])
        check(ccall((:pdcs_get_solution, libpdcs[]), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
            prob[], x, y, s), ctx[])
    finally
        prob[] != C_NULL && ccall((:pdcs_problem_destroy, libpdcs[]), Cint, (Ptr{Cvoid},), prob[])
        ccall((:pdcs_ctx_destroy, libpdcs[]), Cint, (Ptr{Cvoid},), ctx[])
    end
    info = (exit_code = Int(res.exit_code), exit_status = get(EXIT_STATUS, Int(res.exit_code), :unknown),
            pObj = res.pobj, dObj = res.dobj, time = res.time_solve, iter = Int(res.iter), result = res)
    return (x = (recovered_primal = (primal_sol = (x = x,),),),
            y = (recovered_dual = (dual_sol = (y = y,),), slack = (primal_sol = (x = s,),)),
            info = info)
end

end # module
