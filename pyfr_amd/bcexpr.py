"""Compile boundary-value expressions of (x, y, z, t) into the stack
programs ``frhip_bc_eval`` evaluates (opcodes ``FRHIP_BCX_*`` of
``include/frhip.h``).

The reference splices each ``[soln-bcs-*]`` value expression into its
kernel source, with ``ploc`` and a run-time ``t``
(``pyfr/solvers/base/inters.py:54-71``).  The expressions use the
``npeval`` vocabulary (``pyfr/nputil.py:82-117``): numbers, the
``[constants]``, ``x``, ``y``, ``z``, ``t``, ``+ - * /``, unary minus and
the functions below; ``**`` is refused, as there.
"""
import ast

import numpy as np

PUSHC, PUSHV = 0, 1
BINOPS = {ast.Add: 2, ast.Sub: 3, ast.Mult: 4, ast.Div: 5}
FUNCS2 = {'pow': 6, 'min': 7, 'max': 8, 'atan2': 9}
NEG = 10
FUNCS1 = {'sin': 11, 'cos': 12, 'tan': 13, 'exp': 14, 'log': 15, 'sqrt': 16,
          'abs': 17, 'tanh': 18, 'asin': 19, 'acos': 20, 'atan': 21,
          'fabs': 17}
FUNCS2.update({'fmin': 7, 'fmax': 8})
VARS = {'x': 0, 'y': 1, 'z': 2, 't': 3}
MAX_STACK = 16


class ExprError(ValueError):
    pass


def _emit(node, consts, code, pool, varmap=VARS):
    """Post-order emission; returns the stack depth the node needs."""
    def emit(n):
        return _emit(n, consts, code, pool, varmap)
    if isinstance(node, ast.Expression):
        return emit(node.body)
    if isinstance(node, ast.Constant) and isinstance(node.value,
                                                     (int, float)):
        pool.append(float(node.value))
        code.append((PUSHC, len(pool) - 1))
        return 1
    if isinstance(node, ast.Name):
        if node.id in varmap:
            code.append((PUSHV, varmap[node.id]))
        elif node.id == 'pi':
            pool.append(float(np.pi))
            code.append((PUSHC, len(pool) - 1))
        elif node.id in consts:
            pool.append(float(consts[node.id]))
            code.append((PUSHC, len(pool) - 1))
        else:
            raise ExprError(f'Unknown name {node.id!r}')
        return 1
    if isinstance(node, ast.UnaryOp) and isinstance(node.op,
                                                    (ast.USub, ast.UAdd)):
        d = emit(node.operand)
        if isinstance(node.op, ast.USub):
            code.append((NEG, 0))
        return d
    if isinstance(node, ast.BinOp) and type(node.op) in BINOPS:
        da = emit(node.left)
        db = emit(node.right)
        code.append((BINOPS[type(node.op)], 0))
        return max(da, db + 1)
    if (isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and
            not node.keywords):
        fn, args = node.func.id, node.args
        if fn in FUNCS1 and len(args) == 1:
            d = emit(args[0])
            code.append((FUNCS1[fn], 0))
            return d
        if fn in FUNCS2 and len(args) == 2:
            da = emit(args[0])
            db = emit(args[1])
            code.append((FUNCS2[fn], 0))
            return max(da, db + 1)
    raise ExprError(f'Unsupported expression element: {ast.dump(node)}')


def compile_exprs(exprs, consts, params=None, varmap=None):
    """exprs: list of expression strings; consts: {name: value}.  Returns
    (code (ncode, 2) int32, offs (len(exprs) + 1) int32, pool float64).
    params: {name: initial value} of run-time parameters; with it, also
    returns {name: [pool indices]} (set with frhip_bc_program_set_const).
    varmap: {name: variable index} of the program's variables (default x, y,
    z, t of a boundary program)."""
    varmap = VARS if varmap is None else varmap
    code, offs, pool = [], [0], []
    pidx = {}
    if params:
        consts = dict(consts, **params)
    for e in exprs:
        if '**' in e or '^' in e:
            raise ExprError('Direct exponentiation is not supported; use pow')
        try:
            tree = ast.parse(e.strip(), mode='eval')
        except SyntaxError as ex:
            raise ExprError(f'Invalid expression {e!r}') from ex
        n0 = len(code)
        depth = _emit(tree, consts, code, pool, varmap)
        if depth > MAX_STACK:
            raise ExprError(f'Expression too deep: {e!r}')
        if params:
            # the pool entries the parameters' names were emitted as
            pushes = [arg for op, arg in code[n0:] if op == PUSHC]
            for name, idx in zip(_pool_names(tree, consts, varmap), pushes):
                if name in params:
                    pidx.setdefault(name, []).append(idx)
        offs.append(len(code))
    out = (np.array(code, dtype=np.int32).reshape(-1, 2),
           np.array(offs, dtype=np.int32), np.array(pool, dtype=float))
    return out + (pidx,) if params is not None else out


def _pool_names(node, consts, varmap=VARS):
    """For each constant push of the post-order emission, the name it came
    from (None for numbers and pi)."""
    if isinstance(node, ast.Expression):
        return _pool_names(node.body, consts, varmap)
    if isinstance(node, ast.Constant):
        return [None]
    if isinstance(node, ast.Name):
        if node.id in varmap:
            return []
        return [None] if node.id == 'pi' else [node.id]
    if isinstance(node, ast.UnaryOp):
        return _pool_names(node.operand, consts, varmap)
    if isinstance(node, ast.BinOp):
        return (_pool_names(node.left, consts, varmap) +
                _pool_names(node.right, consts, varmap))
    if isinstance(node, ast.Call):
        out = []
        for a in node.args:
            out += _pool_names(a, consts, varmap)
        return out
    return []


def uses_time(expr):
    """True if the expression refers to t."""
    try:
        tree = ast.parse(expr.strip(), mode='eval')
    except SyntaxError:
        return False
    return any(isinstance(n, ast.Name) and n.id == 't'
               for n in ast.walk(tree))
