// Stack programs of (op, arg) pairs (opcodes FRHIP_BCX_* of frhip.h),
// evaluated in double on the device (internal header).  Boundary values and
// the reduction provider's expressions compile to them on the host
// (pyfr_amd/bcexpr.py); the host validates every program before it reaches
// a kernel (operand counts, depth <= EXPR_STACK, index ranges), and the
// guards below keep any access in range regardless.
#pragma once

#include "frhip_common.h"

namespace frhip {

constexpr int EXPR_STACK = 16;

// Program [begin, end) with constants pool[] and variables var[0..nvar).
__device__ __forceinline__ double expr_eval(const int32_t *__restrict__ code,
                                            int begin, int end,
                                            const double *__restrict__ pool,
                                            const double *var, int nvar)
{
    double stk[EXPR_STACK];
    int sp = 0;
    for (int pc = begin; pc < end; ++pc) {
        const int op = code[2*pc], arg = code[2*pc + 1];
        if (op <= FRHIP_BCX_PUSHV) {
            if (sp < EXPR_STACK)
                stk[sp++] = op == FRHIP_BCX_PUSHC
                          ? pool[arg]
                          : var[(arg >= 0 && arg < nvar) ? arg : 0];
            continue;
        }
        if (op <= FRHIP_BCX_ATAN2) {
            if (sp < 2) continue;
            const double b = stk[--sp], a = stk[sp - 1];
            double c;
            switch (op) {
                case FRHIP_BCX_ADD: c = a + b; break;
                case FRHIP_BCX_SUB: c = a - b; break;
                case FRHIP_BCX_MUL: c = a*b; break;
                case FRHIP_BCX_DIV: c = a/b; break;
                case FRHIP_BCX_POW: c = pow(a, b); break;
                case FRHIP_BCX_MIN: c = fmin(a, b); break;
                case FRHIP_BCX_MAX: c = fmax(a, b); break;
                default: c = atan2(a, b); break;
            }
            stk[sp - 1] = c;
            continue;
        }
        if (sp < 1) continue;
        const double a = stk[sp - 1];
        double c;
        switch (op) {
            case FRHIP_BCX_NEG: c = -a; break;
            case FRHIP_BCX_SIN: c = sin(a); break;
            case FRHIP_BCX_COS: c = cos(a); break;
            case FRHIP_BCX_TAN: c = tan(a); break;
            case FRHIP_BCX_EXP: c = exp(a); break;
            case FRHIP_BCX_LOG: c = log(a); break;
            case FRHIP_BCX_SQRT: c = sqrt(a); break;
            case FRHIP_BCX_ABS: c = fabs(a); break;
            case FRHIP_BCX_TANH: c = tanh(a); break;
            case FRHIP_BCX_ASIN: c = asin(a); break;
            case FRHIP_BCX_ACOS: c = acos(a); break;
            default: c = atan(a); break;
        }
        stk[sp - 1] = c;
    }
    return sp > 0 ? stk[0] : 0.0;
}

// Host-side validation of a program of nexprs expressions over nvar
// variables; returns nullptr if valid, else a message.
inline const char *expr_validate(int nexprs, const int32_t *code,
                                 const int32_t *offs, int npool, int nvar)
{
    if (offs[0] != 0) return "program offsets must start at 0";
    for (int r = 0; r < nexprs; ++r) {
        if (offs[r + 1] <= offs[r]) return "empty expression";
        int sp = 0;
        for (int pc = offs[r]; pc < offs[r + 1]; ++pc) {
            const int op = code[2*pc], arg = code[2*pc + 1];
            if (op < FRHIP_BCX_PUSHC || op > FRHIP_BCX_ATAN)
                return "invalid opcode";
            if (op == FRHIP_BCX_PUSHC) {
                if (arg < 0 || arg >= npool)
                    return "constant index out of range";
                ++sp;
            }
            else if (op == FRHIP_BCX_PUSHV) {
                if (arg < 0 || arg >= nvar)
                    return "variable index out of range";
                ++sp;
            }
            else if (op <= FRHIP_BCX_ATAN2) {
                if (sp < 2) return "stack underflow";
                --sp;
            }
            else if (sp < 1)
                return "stack underflow";
            if (sp > EXPR_STACK) return "expression too deep";
        }
        if (sp != 1) return "expression leaves a malformed stack";
    }
    return nullptr;
}

}  // namespace frhip
