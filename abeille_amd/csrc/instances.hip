// instances.hip - device code of the transport kernel variants, one part per
// translation unit: built with -DABX_INST_PART=0..11 (abeille_amd/build.py),
// the parts compile in parallel and link into libabx.so beside abx.hip.
#define ABX_KERNELS_ONLY 1
#include "event.h"

#ifndef ABX_INST_PART
#error "ABX_INST_PART must be defined (0..11)"
#endif

#define ABX_IH(T, TR, SE, SP) ABX_HIST_DECL(, T, TR, SE, SP, false)
#define ABX_IHX(T, TR, SE, SP) ABX_HIST_DECL(, T, TR, SE, SP, true)
#define ABX_IB(T, SE, SP) ABX_BATCH_DECL(, T, SE, SP, false)
#define ABX_IBX(T, SE, SP) ABX_BATCH_DECL(, T, SE, SP, true)
#define ABX_IE(T, TR, SE, SP) ABX_EVENT_DECL(, T, TR, SE, SP)

#if ABX_INST_PART == 0
ABX_VARIANTS_T0(ABX_IH)
#elif ABX_INST_PART == 1
ABX_VARIANTS_T1(ABX_IH)
#elif ABX_INST_PART == 2
ABX_VARIANTS_T2(ABX_IH)
ABX_BATCH_T2(ABX_IB)
#elif ABX_INST_PART == 3
ABX_BATCH_T0(ABX_IB)
#elif ABX_INST_PART == 4
ABX_BATCH_T1(ABX_IB)
ABX_INIT_DECL()
#define ABX_IR(T, SP) ABX_REP_DECL(, T, SP, false)
ABX_REP_ALL(ABX_IR)
#elif ABX_INST_PART == 5
ABX_VARIANTS_T0(ABX_IE)
#elif ABX_INST_PART == 6
ABX_VARIANTS_T1(ABX_IE)
#elif ABX_INST_PART == 7
ABX_VARIANTS_T2(ABX_IE)
// extended-geometry variants (hexagonal lattices, off-axis cylinders)
#elif ABX_INST_PART == 8
ABX_VARIANTS_T0(ABX_IHX)
#elif ABX_INST_PART == 9
ABX_VARIANTS_T1(ABX_IHX)
#elif ABX_INST_PART == 10
ABX_VARIANTS_T2(ABX_IHX)
#define ABX_IRX(T, SP) ABX_REP_DECL(, T, SP, true)
ABX_REP_ALL(ABX_IRX)
#elif ABX_INST_PART == 11
ABX_BATCH_T0(ABX_IBX)
ABX_BATCH_T1(ABX_IBX)
ABX_BATCH_T2(ABX_IBX)
#else
#error "ABX_INST_PART out of range"
#endif
