"""torch.autograd binding of the tape/adjoint path: the derivation as a
differentiable function of the per-axiom parameters.

    inst_0, inst_1, ... = derive_instances(params, engine, terminals)

runs pga_forward_tape with `params` bound as PGA_PT_AXIOM_PARAM values and
returns the instances of each terminal type as (count, 16) float32 tensors
(the 64-byte records: model rows 0-11, then custom, axiom, path, meta).
Backward runs pga_backward with dL/dinstance and returns dL/dparams. The
topology (branches, repetition counts, rule choices) is frozen by the tape
of the forward call; a backward after the engine's tape changed (another
forward, or new axioms, parameters or program) raises instead of replaying
the wrong tape."""
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
        ctx.tape_gen = engine.tape_generation()
        return outs

    @staticmethod
    def backward(ctx, *grads):
        from .engine import PGAError
        from . import abi
        if ctx.engine.tape_generation() != ctx.tape_gen:
            raise PGAError(abi.ERR_NOT_INITIALIZED,
                           "the engine's tape changed since this forward (another forward, or new axioms, "
                           "parameters or program): run the forward again")
        dl = {t: g for t, g in zip(ctx.terminals, grads) if g is not None}
        grad = ctx.engine.backward(dl)
        return grad, None, None


def derive_instances(params, engine, terminals):
    """Instances of `terminals` (a tuple of terminal indices) as a
    differentiable function of `params` ((axioms, params per axiom) float32
    cuda tensor). The engine holds the program, the axioms, a tape
    (`Engine.tape_reserve`) and bound instance buffers."""
    return _Derive.apply(params, engine, tuple(terminals))


class _CenterSizeMSE(torch.autograd.Function):
    @staticmethod
    def forward(ctx, inst, target_centers, target_sizes, wc, ws):
        import ctypes as C

        from .engine import PGAError, load_library
        n = inst.shape[0]
        inst = inst.contiguous()
        tc = target_centers.contiguous()
        ts = target_sizes.contiguous()
        grad = torch.empty_like(inst)
        loss = torch.zeros(1, dtype=torch.float64, device=inst.device)
        st = load_library().pga_center_size_mse(
            C.c_void_p(inst.data_ptr()), C.c_void_p(tc.data_ptr()), C.c_void_p(ts.data_ptr()), n, wc, ws,
            C.c_void_p(grad.data_ptr()), C.c_void_p(loss.data_ptr()),
            C.c_void_p(torch.cuda.current_stream(inst.device).cuda_stream))
        if st != 0:
            raise PGAError(st, "pga_center_size_mse")
        ctx.save_for_backward(grad)
        return loss.to(torch.float32).reshape(())

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        if g.numel() == 1 and float(g) == 1.0:   # loss.backward() on a sum of losses
            return grad, None, None, None, None
        return grad * g, None, None, None, None


def center_size_mse(inst, target_centers, target_sizes, mean=True):
    """The centre/size fitting loss in one kernel (pga_center_size_mse):
    sum over instances of |c - tc|^2 and |s - ts|^2 (c the translation
    column, s the column norms of the 3x3 part), each averaged over the
    instances when `mean`; the gradient rows come out of the same pass.
    Equal to ((c - tc) ** 2).sum(1).mean() + ((s - ts) ** 2).sum(1).mean()
    on centers_and_sizes(inst) up to FP32 rounding."""
    n = inst.shape[0]
    if n == 0:
        return inst.sum() * 0.0
    w = 1.0 / n if mean else 1.0
    return _CenterSizeMSE.apply(inst, target_centers, target_sizes, w, w)


def centers_and_sizes(inst):
    """Centre (translation column) and size (column norms of the 3x3 part)
    of each instance of a (count, 16) record tensor."""
    m = inst[:, :12].reshape(-1, 3, 4)
    return m[:, :, 3], torch.linalg.vector_norm(m[:, :, :3], dim=1)
