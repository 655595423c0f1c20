// The z-pass and column-pass launchers the engine uses, instantiated in
// translation units of their own (pf_pass_inst.hip, one object per group and
// precision) so the device code of the 18 transform lengths compiles in
// parallel; the engine translation units see them as extern templates.
#pragma once

namespace pf {

// single-token names for operators whose template arguments hold commas
template <typename T> using OpPhiRhsE = OpPhiRhs<T, true>;
template <typename T> using OpPhiRhsN = OpPhiRhs<T, false>;
template <typename T> using OpMinRhsE = OpMinRhs<T, true>;
template <typename T> using OpMinRhsN = OpMinRhs<T, false>;
template <typename T> using OpMinEvalE = OpMinEval<T, true>;
template <typename T> using OpMinEvalN = OpMinEval<T, false>;
template <typename T> using OpMinFinishEN = OpMinFinish<T, true, true>;
template <typename T> using OpMinFinishNN = OpMinFinish<T, false, true>;
template <typename T> using OpMinFinishNL = OpMinFinish<T, false, false>;
template <typename T> using OpMinRhsFullE = OpMinRhsFull<T, true>;
template <typename T> using OpMinRhsFullN = OpMinRhsFull<T, false>;
template <typename T> using OpMinEvalFullE = OpMinEvalFull<T, true>;
template <typename T> using OpMinEvalFullN = OpMinEvalFull<T, false>;
template <typename T> using OpMinFinishFullE = OpMinFinishFull<T, true>;
template <typename T> using OpMinFinishFullN = OpMinFinishFull<T, false>;

// (inputs, outputs, operator) per group; the groups balance the compile time
// (groups 0-4, 6 and 7: z-pass operators, 5: the column passes)
#define PF_ZPASS_G0(X, T) X(T, 6, 6, OpHij<T>)
#define PF_ZPASS_G7(X, T) X(T, 6, 6, OpHijInc<T>)
// the solver's z pass at the other LDS budgets the plan tuner tries
// (ZPASS_BUDGETS; index 1 is the operators' default, instantiated above)
#define PF_ZPASS_TUNED_G0(X, T) X(T, 6, 6, OpHij<T>, 0) X(T, 6, 6, OpHij<T>, 2) X(T, 6, 6, OpHij<T>, 3)
#define PF_ZPASS_TUNED_G7(X, T) X(T, 6, 6, OpHijInc<T>, 0) X(T, 6, 6, OpHijInc<T>, 2) X(T, 6, 6, OpHijInc<T>, 3)
#define PF_ZPASS_G1(X, T) X(T, 0, 6, OpElasticCold<T>) X(T, 6, 1, OpPhiRhsE<T>) X(T, 0, 1, OpPhiRhsN<T>)
#define PF_ZPASS_G2(X, T)                                                                          \
  X(T, 1, 1, OpPhiFinish<T>) X(T, 3, 4, OpFlux<T>) X(T, 3, 5, OpFlux<T>) X(T, 1, 0, OpYStar<T>) \
  X(T, 1, 0, OpYFinal<T>)
#define PF_ZPASS_G3(X, T)                                                                               \
  X(T, 6, 1, OpMinRhsE<T>) X(T, 0, 1, OpMinRhsN<T>) X(T, 0, 1, OpMinEvalE<T>) X(T, 0, 1, OpMinEvalN<T>) \
  X(T, 1, 1, OpMinFinishEN<T>) X(T, 1, 1, OpMinFinishNN<T>) X(T, 1, 0, OpMinFinishNL<T>)               \
  X(T, 1, 0, OpMinGfull<T>) X(T, 0, 1, OpMinUpd<T>)
#define PF_ZPASS_G4(X, T)                                                                  \
  X(T, 6, 1, OpMinRhsFullE<T>) X(T, 0, 1, OpMinRhsFullN<T>) X(T, 0, 1, OpMinEvalFullE<T>) \
  X(T, 0, 1, OpMinEvalFullN<T>) X(T, 1, 1, OpMinFinishFullE<T>) X(T, 1, 1, OpMinFinishFullN<T>)
#define PF_ZPASS_G6(X, T)                                                                            \
  X(T, 0, 1, OpMuX<T>) X(T, 0, 1, OpEtaRhs<T>) X(T, 1, 0, OpEtaFinish<T>) X(T, 0, 1, OpLoadReal<T>) \
  X(T, 1, 0, OpStoreReal<T>) X(T, 0, 3, OpLoadF3<T>) X(T, 3, 0, OpStoreF3<T>)
#define PF_ZPASS_LIST(X, T)                                                                             \
  PF_ZPASS_G0(X, T) PF_ZPASS_G1(X, T) PF_ZPASS_G2(X, T) PF_ZPASS_G3(X, T) PF_ZPASS_G4(X, T) PF_ZPASS_G6(X, T) \
  PF_ZPASS_G7(X, T)
#define PF_ZPASS_TUNED_LIST(X, T) PF_ZPASS_TUNED_G0(X, T) PF_ZPASS_TUNED_G7(X, T)
// group 5: the column passes
#define PF_CPASS_LIST(X) X(float, false) X(float, true) X(double, false) X(double, true)

#define PF_ZPASS_EXTERN(T, I, O, OP) \
  extern template int launch_zpass<T, I, O, OP>(int, hipStream_t, const ZPassArgs<T>&, const typename OP::Args&);
#define PF_ZPASS_DEFINE(T, I, O, OP) \
  template int launch_zpass<T, I, O, OP>(int, hipStream_t, const ZPassArgs<T>&, const typename OP::Args&);
#define PF_ZPASS_TUNED_EXTERN(T, I, O, OP, B)                                                                     \
  extern template int launch_zpass<T, I, O, OP, ZPASS_BUDGETS[B]>(int, hipStream_t, const ZPassArgs<T>&, \
                                                                  const typename OP::Args&);
#define PF_ZPASS_TUNED_DEFINE(T, I, O, OP, B)                                                                     \
  template int launch_zpass<T, I, O, OP, ZPASS_BUDGETS[B]>(int, hipStream_t, const ZPassArgs<T>&, \
                                                           const typename OP::Args&);
#define PF_CPASS_EXTERN(T, INV) \
  extern template void launch_cpass<T, INV>(int, hipStream_t, const CPassArgs<T>&, int, int, int);
#define PF_CPASS_DEFINE(T, INV) \
  template void launch_cpass<T, INV>(int, hipStream_t, const CPassArgs<T>&, int, int, int);
#define PF_CAT_(a, b) a##b
#define PF_CAT(a, b) PF_CAT_(a, b)

#ifndef PF_PASS_GROUP
PF_ZPASS_LIST(PF_ZPASS_EXTERN, float)
PF_ZPASS_LIST(PF_ZPASS_EXTERN, double)
PF_ZPASS_TUNED_LIST(PF_ZPASS_TUNED_EXTERN, float)
PF_ZPASS_TUNED_LIST(PF_ZPASS_TUNED_EXTERN, double)
PF_CPASS_LIST(PF_CPASS_EXTERN)
#endif

}  // namespace pf
