"""The boundary-value expression compiler (pyfr_amd.bcexpr): programs run
on a stack machine written here, from the opcode table of include/frhip.h,
against NumPy evaluation of the same expressions; and the expressions it
must refuse."""
import numpy as np
import pytest

from pyfr_amd.bcexpr import ExprError, compile_exprs, uses_time

F1 = {11: np.sin, 12: np.cos, 13: np.tan, 14: np.exp, 15: np.log,
      16: np.sqrt, 17: np.abs, 18: np.tanh, 19: np.arcsin, 20: np.arccos,
      21: np.arctan}
F2 = {2: np.add, 3: np.subtract, 4: np.multiply, 5: np.divide, 6: np.power,
      7: np.minimum, 8: np.maximum, 9: np.arctan2}


def _run(code, offs, pool, x, y, z, t):
    out = []
    var = [x, y, z, t + 0*x]
    for r in range(len(offs) - 1):
        stk = []
        for op, arg in code[offs[r]:offs[r + 1]]:
            if op == 0:
                stk.append(pool[arg] + 0*x)
            elif op == 1:
                stk.append(var[arg])
            elif op in F2:
                b, a = stk.pop(), stk.pop()
                stk.append(F2[op](a, b))
            elif op == 10:
                stk.append(-stk.pop())
            else:
                stk.append(F1[op](stk.pop()))
        assert len(stk) == 1
        out.append(stk[0])
    return out


NPSYMS = {'__builtins__': {}, 'exp': np.exp, 'log': np.log, 'sin': np.sin,
          'cos': np.cos, 'tan': np.tan, 'tanh': np.tanh, 'sqrt': np.sqrt,
          'abs': np.abs, 'pow': np.power, 'pi': np.pi, 'atan': np.arctan,
          'atan2': np.arctan2, 'asin': np.arcsin, 'acos': np.arccos,
          'max': np.maximum, 'min': np.minimum}

EXPRS = [
    '1.0', '-2.5e-3', 'x', 't', '1.0 + 0.05*sin(2.0*pi*t)*cos(z)',
    'pow(1.0 + t, -0.5)*gamma - x/(y + 3.0)', 'min(x, y) - max(z, t)',
    'atan2(y, x + 4.0) + atan(t) + tanh(-x)', 'sqrt(abs(x*y)) + exp(-t*t)',
    'log(2.0 + sin(x)) - asin(0.3*cos(y)) + acos(0.2*t)',
    '-(x - y)*(z + t)/(1.0 + x*x)', 'tan(0.1*x) + +t',
    '((((x + 1.0)*2.0 - y)/3.0 + z)*t - 1.0)*0.5']


def test_programs_match_numpy():
    rng = np.random.default_rng(0)
    x, y, z = rng.uniform(-1, 1, (3, 50))
    consts = {'gamma': 1.4}
    for t in (0.0, 0.37, 2.0):
        code, offs, pool = compile_exprs(EXPRS, consts)
        got = _run(code, offs, pool, x, y, z, t)
        env = dict(consts, x=x, y=y, z=z, t=t)
        for e, g in zip(EXPRS, got):
            ref = eval(e, NPSYMS, env) + 0*x
            assert np.allclose(g, ref, rtol=1e-15, atol=1e-15), e


def test_uses_time():
    assert uses_time('1 + 0.1*t') and uses_time('sin(t)')
    assert not uses_time('1 + tt') and not uses_time('x*y')


@pytest.mark.parametrize('bad', [
    'x**2', 'x^2', 'foo + 1', 'x if t else y', 'x < 1', 'f(x)',
    'sin(x, y)', 'pow(x)', '"s"', 'x.real', '[x]',
    '+'.join(['(x' for _ in range(20)] + [')']*20)])
def test_refused(bad):
    with pytest.raises(ExprError):
        compile_exprs([bad], {})


def test_depth_limit():
    """The left operand waits on the stack while the right one is
    evaluated: left-nested chains need 2 slots, right-nested ones one per
    level, and more than 16 is refused (the kernel's stack)."""
    left = 'x'
    for _ in range(30):
        left = f'({left} + 1.0)'
    code, offs, pool = compile_exprs([left], {})
    assert len(code) == offs[-1] == 61
    right = 'x'
    for _ in range(15):
        right = f'(1.0 + {right})'
    compile_exprs([right], {})              # 16 slots
    with pytest.raises(ExprError):
        compile_exprs([f'(1.0 + {right})'], {})


def _run_vars(code, offs, pool, var):
    """The stack machine over an arbitrary variable list (reduction
    programs)."""
    out = []
    for r in range(len(offs) - 1):
        stk = []
        for op, arg in code[offs[r]:offs[r + 1]]:
            if op == 0:
                stk.append(pool[arg] + 0*var[0])
            elif op == 1:
                stk.append(var[arg])
            elif op in F2:
                b, a = stk.pop(), stk.pop()
                stk.append(F2[op](a, b))
            elif op == 10:
                stk.append(-stk.pop())
            else:
                stk.append(F1[op](stk.pop()))
        assert len(stk) == 1
        out.append(stk[0])
    return out


# expressions the reference's plugins and controllers hand to the reduction
# provider (pyfr/plugins/nancheck.py:21, pyfr/integrators/controllers.py:
# 107-113, pyfr/integrators/implicit/base.py:83-120)
RED_EXPRS = ['x', 'x*x', 'fabs(x)', 'err / (atol + rtol*fabs(curr))',
             '(err / (atol + rtol*fabs(curr)))*(err / (atol + rtol*fabs(curr)))',
             '(x/w)*(x/w)', 'a*b', 'fmax(x, w) - fmin(a, b)']


def test_reduction_programs_match_numpy():
    rng = np.random.default_rng(3)
    names = ['x', 'w', 'a', 'b', 'err', 'curr', 'atol', 'rtol']
    vals = [rng.uniform(0.5, 2.0, 40) for _ in names]
    code, offs, pool = compile_exprs(
        RED_EXPRS, {}, varmap={n: i for i, n in enumerate(names)})
    got = _run_vars(code, offs, pool, vals)
    env = dict(zip(names, vals), fabs=np.abs, fmax=np.maximum,
               fmin=np.minimum)
    for e, g in zip(RED_EXPRS, got):
        ref = eval(e, {'__builtins__': {}}, env)
        assert np.allclose(g, ref, rtol=1e-15, atol=0), e


@pytest.mark.parametrize('op', [-1, -7, 22, 1000])
def test_library_refuses_invalid_opcodes(op):
    """Both program constructors reject opcodes outside the table before
    any device call (so this runs without a GPU)."""
    import ctypes as ct

    from pyfr_amd._lib import FrHip, FrHipError
    lib = FrHip()
    # push x, push x, <op>: valid as a binary op if op were in range
    code = np.array([[1, 0], [1, 0], [op, 0]], dtype=np.int32)
    offs = np.array([0, 3], dtype=np.int32)
    pool = np.zeros(1)
    p = ct.c_void_p()
    with pytest.raises(FrHipError, match='invalid opcode'):
        lib.bc_program_create(ct.byref(p), 1, code.ctypes.data,
                              offs.ctypes.data, pool.ctypes.data, 1)
    with pytest.raises(FrHipError, match='invalid opcode'):
        lib.expr_program_create(ct.byref(p), 1, 2, code.ctypes.data,
                                offs.ctypes.data, pool.ctypes.data, 1)


def test_library_refuses_bad_variable_index():
    import ctypes as ct

    from pyfr_amd._lib import FrHip, FrHipError
    lib = FrHip()
    code = np.array([[1, 3]], dtype=np.int32)      # variable 3 of 3
    offs = np.array([0, 1], dtype=np.int32)
    p = ct.c_void_p()
    with pytest.raises(FrHipError, match='variable index'):
        lib.expr_program_create(ct.byref(p), 1, 3, code.ctypes.data,
                                offs.ctypes.data, None, 0)
