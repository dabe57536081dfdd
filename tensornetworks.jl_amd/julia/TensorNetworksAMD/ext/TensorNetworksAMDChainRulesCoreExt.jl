# ChainRulesCore rule for the coupling-parametrised energy E(c) = <psi|H(c)|psi> of
# TensorNetworksAMD.coupling_energy, loaded when the caller has ChainRulesCore available
# (package extension). The reference has no differentiation; the pullback is the device
# backward pass of the environment sweep (tnh_mpo_expect_grad, checkpointed every `block`
# sites), mapped to the couplings through the automaton slots of each term.
module TensorNetworksAMDChainRulesCoreExt

using ChainRulesCore
using TensorNetworks: GMPS, OpList, Sitetypes
import TensorNetworksAMD: coupling_energy, coupling_energy_grad, default_context

function ChainRulesCore.rrule(::typeof(coupling_energy), psi::GMPS, st::Sitetypes, ol::OpList,
                              c::AbstractVector; block::Int = 8, ctx = default_context())
    E, g = coupling_energy_grad(psi, st, ol, c; block = block, ctx = ctx)
    function coupling_energy_pullback(Ebar)
        return (NoTangent(), NoTangent(), NoTangent(), NoTangent(), unthunk(Ebar) .* g)
    end
    return E, coupling_energy_pullback
end

end # module
