"""The FP64 p-form of a PCG iteration (HOBAK_TUNE_P_FORM): the previous iteration is finished and
p_new = z + beta p_old formed once per entry by its own kernel, then the SpMV gathers p_new alone.
Every value is computed by the same operations in the same order as in the two-kernel form, so
the bar is bitwise equality of positions, velocities and PCG iteration counts."""
import numpy as np
import pytest

from hobakv1_amd import lib, scenes

pytestmark = pytest.mark.gpu


def _steps(ctx, scs, steps, batched):
    insts = [lib.Instance(ctx, sc) for sc in scs]
    its = [[] for _ in scs]
    for _ in range(steps):
        if batched:
            for i, st in enumerate(lib.step_batch(ctx, insts)):
                its[i].append(st["pcg_iters"])
        else:
            for i, inst in enumerate(insts):
                its[i].append(inst.step()["pcg_iters"])
    out = [(inst.state(), its[i]) for i, inst in enumerate(insts)]
    for inst in insts:
        inst.close()
    return out


@pytest.mark.parametrize("precond", [scenes.PRECOND_JACOBI, scenes.PRECOND_BLOCK_JACOBI])
def test_pform_equals_two_kernel_form_c1(precond):
    sc = scenes.config1()
    res = []
    for pf in (0, 2):
        ctx = lib.Context(0)
        ctx.set_tuning(p_form=pf)
        inst = lib.Instance(ctx, sc, precond=precond)
        its = [inst.step()["pcg_iters"] for _ in range(20)]
        res.append((inst.state(), its))
        inst.close()
        ctx.close()
    (ua, va), ia = res[0]
    (ub, vb), ib = res[1]
    assert ia == ib
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)


def test_pform_equals_two_kernel_form_batch_with_contact():
    """A batch of mixed instances (sizes, materials, a contact drop scene, BDF_1 off) stepped
    together in both forms."""
    scs = scenes.batch_instances(count=4, seed=31, lo=3, hi=8) + [scenes.drop_scene()]
    res = []
    for pf in (0, 2):
        ctx = lib.Context(0)
        ctx.set_tuning(p_form=pf)
        res.append(_steps(ctx, scs, 5, batched=True))
        ctx.close()
    for ((ua, va), ia), ((ub, vb), ib) in zip(res[0], res[1]):
        assert ia == ib
        assert np.array_equal(ua, ub) and np.array_equal(va, vb)


def test_pform_refuses_bad_key(gpu_ctx):
    with pytest.raises(lib.HobakError):
        gpu_ctx.set_tuning(p_form=3)
