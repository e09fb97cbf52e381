"""torch.autograd binding of the tape/adjoint path: the derivation as a
differentiable function of the per-axiom parameters.

    inst_0, inst_1, ... = derive_instances(params, engine, terminals)

runs pga_forward_tape with `params` bound as PGA_PT_AXIOM_PARAM values and
returns the instances of each terminal type as (count, 16) float32 tensors
(the 64-byte records: model rows 0-11, then custom, axiom, path, meta).
Backward runs pga_backward with dL/dinstance and returns dL/dparams. The
topology (branches, repetition counts, rule choices) is frozen by the tape
of the forward call."""
import torch


class _Derive(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, engine, terminals):
        engine.set_axiom_params(params)
        engine.forward_tape()
        counts = engine.counts(max(terminals) + 1)
        outs = tuple(engine.buffer(t)[: int(counts[t])].clone() for t in terminals)
        ctx.engine = engine
        ctx.terminals = terminals
        return outs

    @staticmethod
    def backward(ctx, *grads):
        dl = {t: g for t, g in zip(ctx.terminals, grads) if g is not None}
        grad = ctx.engine.backward(dl)
        return grad, None, None


def derive_instances(params, engine, terminals):
    """Instances of `terminals` (a tuple of terminal indices) as a
    differentiable function of `params` ((axioms, params per axiom) float32
    cuda tensor). The engine holds the program, the axioms, a tape
    (`Engine.tape_reserve`) and bound instance buffers."""
    return _Derive.apply(params, engine, tuple(terminals))


def centers_and_sizes(inst):
    """Centre (translation column) and size (column norms of the 3x3 part)
    of each instance of a (count, 16) record tensor."""
    m = inst[:, :12].reshape(-1, 3, 4)
    return m[:, :, 3], torch.linalg.vector_norm(m[:, :, :3], dim=1)
