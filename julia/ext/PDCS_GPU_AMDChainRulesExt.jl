"""
Reverse-mode rule of `PDCS_GPU_AMD.conic_layer` for ChainRulesCore-based AD
(Zygote and others): the pullback of x with respect to (c, G, h) is the GPU
adjoint of the KKT fixed point (`pdcs_adjoint`).
"""
module PDCS_GPU_AMDChainRulesExt

using ChainRulesCore
using PDCS_GPU_AMD

function ChainRulesCore.rrule(::typeof(PDCS_GPU_AMD.conic_layer), c, G, h; kw...)
    x, R = PDCS_GPU_AMD.conic_layer_forward(c, G, h; kw...)
    function conic_layer_pullback(x̄)
        dx = unthunk(x̄)
        dx isa AbstractZero && return NoTangent(), ZeroTangent(), ZeroTangent(), ZeroTangent()
        dc, dG, dh = PDCS_GPU_AMD.conic_layer_adjoint(R, dx)
        return NoTangent(), dc, dG, dh
    end
    return x, conic_layer_pullback
end

end # module
