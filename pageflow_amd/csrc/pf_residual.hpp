// The residual program of a join (PF_RES_* in pageflow_join.h): its
// kernel-argument form, the comparison rules, host validation, and the
// per-pair interpreter the partitioned join evaluates inside its count and
// write passes.  The standalone pair filter (join_core.hip) and the
// partitioned join (hash_join.hip) share it.
//
// Comparison semantics follow operators/operator_utils.py:438-478 of the
// reference: integers compare in int64, float64 comparisons in float64 with
// integer columns converted, `=` on a float64 left operand is isclose.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "pf_common.hpp"
#include "../../include/pageflow_join.h"

namespace pf {

struct RfProgram {
  const void* col[PF_RES_MAX_COLS];
  int64_t rows[PF_RES_MAX_COLS];
  int32_t type[PF_RES_MAX_COLS];
  int32_t side[PF_RES_MAX_COLS];
  int64_t ins[PF_RES_MAX_INS][3];
  int32_t n_ins;
};

// torch.isclose(a, b, rtol=1e-5, atol=1e-5) in float64: a == b, or a
// finite |a - b| within atol + |rtol * b|, each step rounded once.
__device__ __forceinline__ bool rf_close(double a, double b) {
  if (a == b) return true;
  const double err = fabs(__dsub_rn(a, b));
  const double tol = __dadd_rn(1e-5, fabs(__dmul_rn(1e-5, b)));
  return isfinite(err) && err <= tol;
}

__device__ __forceinline__ bool rf_rel_i(int rel, int64_t a, int64_t b) {
  switch (rel) {
    case PF_REL_EQ: return a == b;
    case PF_REL_NE: return a != b;
    case PF_REL_LT: return a < b;
    case PF_REL_LE: return a <= b;
    case PF_REL_GT: return a > b;
    default: return a >= b;
  }
}

__device__ __forceinline__ bool rf_rel_f(int rel, double a, double b) {
  switch (rel) {
    case PF_REL_EQ: return a == b;
    case PF_REL_NE: return a != b;
    case PF_REL_LT: return a < b;
    case PF_REL_LE: return a <= b;
    case PF_REL_GT: return a > b;
    case PF_REL_GE: return a >= b;
    default: return rf_close(a, b);
  }
}

// A column value widened to 64 bits: integers sign-extended (uint8
// zero-extended), a float64 as its bits.
__device__ __forceinline__ int64_t rf_widen(const void* p, int type, int64_t i) {
  switch (type) {
    case PF_COL_I32: return (int64_t)((const int32_t*)p)[i];
    case PF_COL_I16: return (int64_t)((const int16_t*)p)[i];
    case PF_COL_I8: return (int64_t)((const int8_t*)p)[i];
    case PF_COL_U8: return (int64_t)((const uint8_t*)p)[i];
    default: return ((const int64_t*)p)[i];  // int64, or a float64's bits
  }
}

// Checks the program: opcodes, relations, slots, operand domains and the
// stack depth (at most 32, exactly one result).
inline int rf_validate(const RfProgram& P, int n_cols) {
  int depth = 0;
  for (int k = 0; k < P.n_ins; ++k) {
    const int64_t w0 = P.ins[k][0];
    const int opc = (int)(w0 & 0xff);
    const bool f64 = (w0 & PF_RES_F64) != 0;
    if (opc == PF_RES_CMP) {
      const int rel = (int)((w0 >> 8) & 0xff);
      if (rel < PF_REL_EQ || rel > PF_REL_CLOSE || (rel == PF_REL_CLOSE && !f64))
        return fail("pair_filter: instruction " + std::to_string(k) + " has an unknown relation");
      for (int o = 0; o < 2; ++o) {
        if (!(w0 & (o ? PF_RES_B_COL : PF_RES_A_COL))) continue;
        const int64_t s = P.ins[k][1 + o];
        if (s < 0 || s >= n_cols)
          return fail("pair_filter: instruction " + std::to_string(k) + " names column slot " +
                      std::to_string(s) + " of " + std::to_string(n_cols));
        if (P.type[s] == PF_COL_F64 && !f64)
          return fail("pair_filter: a float64 column compares in float64 only");
      }
      ++depth;
    } else if (opc == PF_RES_TRUE) {
      ++depth;
    } else if (opc == PF_RES_AND || opc == PF_RES_OR) {
      if (depth < 2) return fail("pair_filter: AND/OR on fewer than two results");
      --depth;
    } else if (opc == PF_RES_NOT) {
      if (depth < 1) return fail("pair_filter: NOT on no result");
    } else {
      return fail("pair_filter: instruction " + std::to_string(k) + " has opcode " +
                  std::to_string(opc));
    }
    if (depth > 32) return fail("pair_filter: the program is deeper than 32 results");
  }
  if (depth != 1) return fail("pair_filter: the program leaves " + std::to_string(depth) +
                              " results, not one");
  return 0;
}

// The C ABI's program arrays as an RfProgram, checked (column types, row
// pointers, then the program itself).
inline int rf_program_of(const void* const* cols, const int32_t* col_types,
                         const int32_t* col_sides, const int64_t* col_rows, int32_t n_cols,
                         const int64_t* ins, int32_t n_ins, RfProgram* P) {
  if (n_cols < 0 || n_cols > PF_RES_MAX_COLS)
    return fail("pair_filter: at most " + std::to_string(PF_RES_MAX_COLS) + " columns");
  if (n_ins < 1 || n_ins > PF_RES_MAX_INS)
    return fail("pair_filter: 1 to " + std::to_string(PF_RES_MAX_INS) + " instructions");
  if (!ins || (n_cols > 0 && (!cols || !col_types || !col_sides || !col_rows)))
    return fail("pair_filter: missing program or column arrays");
  *P = RfProgram{};
  for (int s = 0; s < n_cols; ++s) {
    if (col_types[s] < PF_COL_I64 || col_types[s] > PF_COL_F64)
      return fail("pair_filter: column slot " + std::to_string(s) + " has an unknown type");
    if (col_rows[s] < 0 || (col_rows[s] > 0 && !cols[s]))
      return fail("pair_filter: column slot " + std::to_string(s) + " has no rows pointer");
    P->col[s] = cols[s];
    P->rows[s] = col_rows[s];
    P->type[s] = col_types[s];
    P->side[s] = col_sides[s] != 0;
  }
  P->n_ins = n_ins;
  for (int k = 0; k < n_ins; ++k)
    for (int q = 0; q < 3; ++q) P->ins[k][q] = ins[3 * k + q];
  return rf_validate(*P, n_cols);
}

// ------------------------------------------- the join's pair interpreter --
//
// Inside the partitioned join a column slot's value comes from one of:
//   RS_ROW    the column itself, at the pair's build or probe row (any
//             PF_COL_* type);
//   RS_BLDS   the first build column, gathered into partition order and
//             staged in LDS with the build chunk (index: the chunk row);
//   RS_BPART  the second build column in partition order in HBM (index:
//             the chunk's first row + the chunk row);
//   RS_P0/P1  a probe column carried through the probe side's partition
//             (the probe row's value, held with its hash).
// Every source but RS_ROW holds the widened value (rf_widen).
enum : int32_t { RS_ROW = 0, RS_BLDS = 1, RS_BPART = 2, RS_P0 = 3, RS_P1 = 4 };

struct JoinResidual {
  RfProgram P;
  int32_t src[PF_RES_MAX_COLS];
  const int64_t* bpart1;  // the RS_BPART column in partition order
};

// The operands of one pair.
struct PairCtx {
  const int64_t* bv;    // RS_BLDS values of the chunk (LDS)
  const int32_t* rows;  // the chunk's build row ids (LDS)
  int li;               // the build row's index in its chunk
  int64_t cb;           // the chunk's first row in partition order
  int64_t prow;         // the probe row
  int64_t pv0, pv1;     // its carried values
};

__device__ __forceinline__ int64_t jr_operand(const JoinResidual& R, int64_t word, bool is_col,
                                              bool f64, const PairCtx& c) {
  if (!is_col) return word;
  const int s = (int)word;
  int64_t v;
  switch (R.src[s]) {
    case RS_BLDS: v = c.bv[c.li]; break;
    case RS_BPART: v = R.bpart1[c.cb + c.li]; break;
    case RS_P0: v = c.pv0; break;
    case RS_P1: v = c.pv1; break;
    default:
      v = rf_widen(R.P.col[s], R.P.type[s], R.P.side[s] ? c.prow : (int64_t)c.rows[c.li]);
      break;
  }
  if (f64 && R.P.type[s] != PF_COL_F64) v = __double_as_longlong((double)v);
  return v;
}

// The program's result on one pair (the instruction words are uniform
// across the wave: kernel arguments).
__device__ __forceinline__ bool jr_eval(const JoinResidual& R, const PairCtx& c) {
  uint32_t stk = 0;
  for (int pc = 0; pc < R.P.n_ins; ++pc) {
    const int64_t w0 = R.P.ins[pc][0];
    const int opc = (int)(w0 & 0xff);
    if (opc == PF_RES_CMP) {
      const int rel = (int)((w0 >> 8) & 0xff);
      const bool f64 = (w0 & PF_RES_F64) != 0;
      const int64_t a = jr_operand(R, R.P.ins[pc][1], (w0 & PF_RES_A_COL) != 0, f64, c);
      const int64_t b = jr_operand(R, R.P.ins[pc][2], (w0 & PF_RES_B_COL) != 0, f64, c);
      const bool r = f64 ? rf_rel_f(rel, __longlong_as_double(a), __longlong_as_double(b))
                         : rf_rel_i(rel, a, b);
      stk = (stk << 1) | (uint32_t)r;
    } else if (opc == PF_RES_AND) {
      const uint32_t t = stk & 1u;
      stk >>= 1;
      stk &= ~1u | t;
    } else if (opc == PF_RES_OR) {
      const uint32_t t = stk & 1u;
      stk = (stk >> 1) | t;
    } else if (opc == PF_RES_NOT) {
      stk ^= 1u;
    } else {  // PF_RES_TRUE
      stk = (stk << 1) | 1u;
    }
  }
  return (stk & 1u) != 0;
}

}  // namespace pf
