/* Drop-in C++ surface of the reference GPU apply, same names and signatures
 * as the reference's include/BEM3dLaplaceGPU.h:8-18, implemented over the
 * C-ABI in easybem_hip.h (easybem_amd/csrc/legacy_api.cpp).
 *
 * Like the reference, one mesh is live at a time (module state between
 * initGPU and closeGPU).  Differences: nblocks/nthreads are accepted and
 * ignored (launch shapes come from the library), nothing is printed, the
 * apply runs in FP64 (near field and far field) unless setGPUFarPrecision(1)
 * selects the FP32 far field, and misuse (no initGPU, rejected mesh,
 * unsupported order) prints the library's message to stderr and aborts
 * instead of reading uninitialised memory.  applyGPULaplace3dHyperSing, an
 * empty stub in the reference (src/BEM3dLaplaceGPU.cu:1554-1555), applies
 * D = sum_c C_c^T V C_c with the surface curls C_c of getLaplace3dCurl, the
 * formula getLaplace3d assembles densely (src/BEM3dLaplace.cpp:650-654,
 * 767-771).
 */
#ifndef BEM3DLAPLACEGPU_H
#define BEM3DLAPLACEGPU_H

void initGPU(int nNodes, const double *nodes, int nElements,
             const int *elements, int order, int nblocks, int nthreads);

void closeGPU();

void applyGPULaplace3d1Layer(const double *t, double *v, int ansatz = 0);

void applyGPULaplace3d2Layer(const double *u, double *q);

void applyGPULaplace3dHyperSing(const double *u, double *q);

/* Extension: 0 = FP64 far field (default), 1 = FP32 far field. */
void setGPUFarPrecision(int fp32);

/* Extension: sizes of the live mesh (the reference binding reads the
 * globals ne / nnodes, python/bindings.cpp:5). */
int gpuNumElements();
int gpuNumNodes();

#endif
