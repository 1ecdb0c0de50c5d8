// instances.h - the compiled variants of the transport kernels.
//
// Every (tracking, trace, secondary stack, precision, LDS staging,
// extended geometry) variant of the big transport kernels is instantiated
// once, in one of the parts of
// instances.hip (each part is its own translation unit, compiled in
// parallel).  Everywhere else the variants are `extern`, so the host code of
// abx.hip launches them without compiling their device code again.
//
// The lists follow dispatch_variant (event.h): traces only in FP64,
// carter tracking always with the secondary stack.
#pragma once

// (TRK, TRACE, SEC, SPLIT) of the history and event kernels
#define ABX_VARIANTS_T0(X)                                                                    \
  X(0, false, false, false) X(0, false, true, false) X(0, true, false, false)                 \
  X(0, true, true, false) X(0, false, false, true) X(0, false, true, true)
#define ABX_VARIANTS_T1(X)                                                                    \
  X(1, false, false, false) X(1, false, true, false) X(1, true, false, false)                 \
  X(1, true, true, false) X(1, false, false, true) X(1, false, true, true)
#define ABX_VARIANTS_T2(X) X(2, false, true, false) X(2, true, true, false) X(2, false, true, true)

// (TRK, SEC, SPLIT) of the batch kernel (never traced)
#define ABX_BATCH_T0(X) X(0, false, false) X(0, true, false) X(0, false, true) X(0, true, true)
#define ABX_BATCH_T1(X) X(1, false, false) X(1, true, false) X(1, false, true) X(1, true, true)
#define ABX_BATCH_T2(X) X(2, true, false) X(2, true, true)

#define ABX_HIST_DECL(EXT, T, TR, SE, SP, EX)                                                 \
  EXT template __global__ void k_history<T, TR, SE, SP, true, false, EX>(DevProblem, KCtx,    \
                                                                         unsigned long long*); \
  EXT template __global__ void k_history<T, TR, SE, SP, false, false, EX>(DevProblem, KCtx,   \
                                                                          unsigned long long*);
// replica mode: k-eigenvalue chains (no trace, no secondary stack)
#define ABX_REP_DECL(EXT, T, SP, EX)                                                          \
  EXT template __global__ void k_history<T, false, false, SP, true, true, EX>(                \
      DevProblem, KCtx, unsigned long long*);                                                 \
  EXT template __global__ void k_history<T, false, false, SP, false, true, EX>(               \
      DevProblem, KCtx, unsigned long long*);
#define ABX_REP_ALL(X) X(0, false) X(0, true) X(1, false) X(1, true)
#define ABX_BATCH_DECL(EXT, T, SE, SP, EX)                                                    \
  EXT template __global__ void k_batch_chunks<T, SE, SP, true, EX>(                           \
      const DevProblem*, const KCtx*, const BatchChunk*, int, int, int, unsigned*);           \
  EXT template __global__ void k_batch_chunks<T, SE, SP, false, EX>(                          \
      const DevProblem*, const KCtx*, const BatchChunk*, int, int, int, unsigned*);
#define ABX_EVENT_DECL(EXT, T, TR, SE, SP)                                                    \
  EXT template __global__ void k_ev_flight<T, TR, SE, SP>(DevProblem, KCtx, EvState,          \
                                                          const int32_t*, const unsigned*,    \
                                                          int32_t*, unsigned*, int32_t*,      \
                                                          unsigned*);                         \
  EXT template __global__ void k_ev_collide<T, TR, SE, SP>(DevProblem, KCtx, EvState,         \
                                                           const int32_t*, const unsigned*,   \
                                                           int32_t*, unsigned*);              \
  EXT template __global__ void k_ev_tail<T, TR, SE, SP>(DevProblem, KCtx, EvState,            \
                                                        const int32_t*, const unsigned*,      \
                                                        const int32_t*, const unsigned*);
#define ABX_INIT_DECL(EXT)                                                                    \
  EXT template __global__ void k_ev_init<false, false>(DevProblem, KCtx, EvState, int32_t*,   \
                                                       unsigned*);                            \
  EXT template __global__ void k_ev_init<false, true>(DevProblem, KCtx, EvState, int32_t*,    \
                                                      unsigned*);                             \
  EXT template __global__ void k_ev_init<true, false>(DevProblem, KCtx, EvState, int32_t*,    \
                                                      unsigned*);                             \
  EXT template __global__ void k_ev_init<true, true>(DevProblem, KCtx, EvState, int32_t*,     \
                                                     unsigned*);

#define ABX_XH(T, TR, SE, SP)                                                                 \
  ABX_HIST_DECL(extern, T, TR, SE, SP, false) ABX_HIST_DECL(extern, T, TR, SE, SP, true)
#define ABX_XB(T, SE, SP)                                                                     \
  ABX_BATCH_DECL(extern, T, SE, SP, false) ABX_BATCH_DECL(extern, T, SE, SP, true)
#define ABX_XE(T, TR, SE, SP) ABX_EVENT_DECL(extern, T, TR, SE, SP)
ABX_VARIANTS_T0(ABX_XH)
ABX_VARIANTS_T1(ABX_XH)
ABX_VARIANTS_T2(ABX_XH)
ABX_BATCH_T0(ABX_XB)
ABX_BATCH_T1(ABX_XB)
ABX_BATCH_T2(ABX_XB)
ABX_VARIANTS_T0(ABX_XE)
ABX_VARIANTS_T1(ABX_XE)
ABX_VARIANTS_T2(ABX_XE)
ABX_INIT_DECL(extern)
#define ABX_XR(T, SP) ABX_REP_DECL(extern, T, SP, false) ABX_REP_DECL(extern, T, SP, true)
ABX_REP_ALL(ABX_XR)
#undef ABX_XR
#undef ABX_XH
#undef ABX_XB
#undef ABX_XE
