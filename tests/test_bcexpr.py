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
