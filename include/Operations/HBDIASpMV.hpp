// HBDIASpMV.hpp -- the HBDIA sparse matrix-vector product y = A x.
//
// Drop-in for the reference's include/Operations/HBDIASpMV.hpp:8-21.
//   hbdiaSpMV(A, x, y, execCOOCPU, execCOOGPU)
//     * no COO fallback: one blocked-DIA launch on A's BDIA stream;
//     * execCOOCPU (wins when both flags are set): the reference sequence --
//       BDIA launch, x copied to the host, OpenMP COO SpMV into the vector's
//       COO buffer, copy back, y += coo on the ADD stream after both events;
//     * execCOOGPU: one fused launch (blocked-DIA + COO + merge) on the BDIA
//       stream; needs ExecutionMode::GPU_COO at conversion time;
//     * COO present and both flags false: the BDIA part is launched, then
//       std::runtime_error("Invalid COO execution flags") is thrown.
//   Asynchronous like the reference: call A.synchronize() (or
//   hbdia_device_synchronize()) before reading y.  Returns true.
#ifndef HBDIASPMV_HPP
#define HBDIASPMV_HPP

#include "../Format/HBDIA.hpp"
#include "../Format/HBDIAVector.hpp"

template <typename T>
bool hbdiaSpMV(HBDIA<T>& matrix, const HBDIAVector<T>& inputVector, HBDIAVector<T>& outputVector,
               bool execCooCPU, bool execCooGPU);

// Multiple right-hand sides (no reference counterpart): for j < nrhs,
// Y[:, j] = A X[:, j] with device arrays X and Y in column-major order
// (leading dimensions ldx, ldy >= numRows).  One launch on the BDIA stream;
// with execCOOGPU the fallback is fused as in hbdiaSpMV's GPU branch.  Each
// column equals hbdiaSpMV on that column bit for bit.  Needs a block width that
// is a multiple of 64.  Asynchronous.
template <typename T>
bool hbdiaSpMVMulti(HBDIA<T>& matrix, const T* X, int64_t ldx, T* Y, int64_t ldy, int nrhs, bool execCOOGPU);

// The same with X and Y row-major (interleaved): row r's nrhs values are
// adjacent, X[r * nrhs + j].  A row's values must be a multiple of 16 bytes
// and at most 512 (FP64: nrhs even and <= 64; FP32: nrhs a multiple of 4 and
// <= 128); nrhs == 1 takes hbdiaSpMV's kernel.  X holds max(numRows, numCols)
// rows when the matrix has a fallback, Y numRows rows.  Each column equals
// hbdiaSpMV on that column bit for bit.  Asynchronous.
template <typename T>
bool hbdiaSpMVMultiRowMajor(HBDIA<T>& matrix, const T* X, T* Y, int nrhs, bool execCOOGPU);

// COO fallback on the host (reference HBDIASpMV.cpp:101-131).
template <typename T>
void executeCOOOnCPU(HBDIA<T>& matrix, const HBDIAVector<T>& inputVector,
                     HBDIAVector<T>& outputVector, int numRows, hbdia_stream_t sD2H,
                     hbdia_stream_t sH2D, hbdia_stream_t sCOO, hbdia_event_t cooEvent);

// COO fallback on the device into the vector's COO buffer (reference
// HBDIASpMV.cpp:133-183, which calls cuSPARSE there).
template <typename T>
void executeCOOOnGPU(HBDIA<T>& matrix, const HBDIAVector<T>& inputVector,
                     HBDIAVector<T>& outputVector, int numRows, hbdia_stream_t sCOO,
                     hbdia_event_t cooEvent);

// OpenMP COO kernel of the CPU path (reference HBDIASpMV.cpp:35-99): rows with
// fallback entries get their sequential sum in entry order; others are left.
template <typename T>
void hbdia_cpu_coo_spmv(const std::vector<int>& cpuRowIndices, const std::vector<int>& cpuColIndices,
                        const std::vector<T>& cpuValues, const T* inputVector, int numRows,
                        T* cpuResults);

#endif  // HBDIASPMV_HPP
