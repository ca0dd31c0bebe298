// HBDIAStats.cpp -- storage-cost report (reference benchmarking/benchmark.cpp:144-230).
#include "Operations/HBDIAStats.hpp"

#include <algorithm>
#include <sstream>
#include <stdexcept>

template <typename T>
HBDIAStorageReport hbdiaStorageReport(const HBDIA<T>& m) {
  if (!m.isHBDIAFormat()) throw std::invalid_argument("storage report needs a matrix in HBDIA format");
  HBDIAStorageReport r;
  const int64_t vb = static_cast<int64_t>(sizeof(T));
  r.valueBytes = vb;
  r.rows = m.getNumRows();
  r.cols = m.getNumCols();
  r.nnz = m.getNumNonZeros();
  r.numberDiagonals = m.getNumberDiagonals();
  r.numBlocks = m.getNumBlocks();
  r.blockWidth = m.getBlockWidth();
  r.threshold = m.getThreshold();
  r.maxCooEntries = m.getMaxCooEntries();
  r.cooFallbackEntries = static_cast<int64_t>(m.getCpuValues().size());
  r.hbdiaDataSize = static_cast<int64_t>(m.getNumStoredValues());

  const auto& hb = m.getHistogramBlocks();
  const auto& hn = m.getHistogramNnz();
  for (size_t i = 0; i < hb.size(); ++i) r.storedOffsets += static_cast<int64_t>(i) * hb[i];
  for (size_t i = 0; i < hb.size(); ++i) {
    const int64_t nz = i < hn.size() ? hn[i] : 0;
    if (hb[i] > 0 || nz > 0) r.histogram.push_back({static_cast<int64_t>(i), hb[i], nz});
  }

  // :181-186 data, block pointers, one offset per (block, diagonal), fallback COO
  r.hbdiaBytes = r.hbdiaDataSize * vb + (r.numBlocks + 1) * 4 + r.storedOffsets * 4 + r.cooFallbackEntries * 8 +
                 r.cooFallbackEntries * vb;
  // :189-190
  r.cooBytes = r.nnz * 8 + r.nnz * vb;
  // :193-194 one full-length diagonal per distinct offset
  r.diaBytes = r.numberDiagonals * std::min(r.rows, r.cols) * vb + r.numberDiagonals * 4;
  // :197-215 slices of blockWidth rows padded to the slice's longest row, plus slice pointers
  {
    const auto& per = m.getNnzPerRow();
    const int64_t bw = r.blockWidth;
    int64_t mx = 0, sell = 0;
    for (size_t i = 0; i < per.size(); ++i) {
      if (static_cast<int64_t>(i) % bw == 0 && i > 0) {
        sell += mx * bw * vb + mx * bw * 4;
        mx = 0;
      }
      mx = std::max<int64_t>(mx, per[i]);
    }
    sell += mx * bw * vb + mx * bw * 4;
    sell += ((r.rows + bw - 1) / bw) * 4;
    r.sellBytes = sell;
  }
  // :218-220
  r.csrBytes = r.nnz * vb + r.nnz * 4 + (r.rows + 1) * 4;
  const double coo = static_cast<double>(r.cooBytes);
  r.hbdiaVsCoo = r.hbdiaBytes / coo;
  r.diaVsCoo = r.diaBytes / coo;
  r.csrVsCoo = r.csrBytes / coo;
  r.sellVsCoo = r.sellBytes / coo;
  r.sellApplies = r.threshold == 1;
  return r;
}

std::string hbdiaStorageReportCSV(const HBDIAStorageReport& r, const std::string& name) {
  std::ostringstream o;
  o << "name,rows,cols,nnz,num_diagonals,num_blocks,block_width,threshold,max_coo_entries,"
       "coo_fallback_entries,hbdia_data_size,hbdia_cost_bytes,coo_cost_bytes,dia_cost_bytes,sell_cost_bytes,"
       "csr_cost_bytes,hbdia_vs_coo,dia_vs_coo,csr_vs_coo,sell_vs_coo\n";
  o << name << ',' << r.rows << ',' << r.cols << ',' << r.nnz << ',' << r.numberDiagonals << ',' << r.numBlocks
    << ',' << r.blockWidth << ',' << r.threshold << ',' << r.maxCooEntries << ',' << r.cooFallbackEntries << ','
    << r.hbdiaDataSize << ',' << r.hbdiaBytes << ',' << r.cooBytes << ',' << r.diaBytes << ',' << r.sellBytes
    << ',' << r.csrBytes << ',' << std::to_string(r.hbdiaVsCoo) << ',' << std::to_string(r.diaVsCoo) << ','
    << std::to_string(r.csrVsCoo) << ',' << std::to_string(r.sellVsCoo) << "\n\n";
  o << "diagonal_count,blocks_with_diagonals,nnz_in_blocks\n";
  for (const auto& h : r.histogram) o << h.diagonals << ',' << h.blocks << ',' << h.nnz << '\n';
  return o.str();
}

template HBDIAStorageReport hbdiaStorageReport<double>(const HBDIA<double>&);
template HBDIAStorageReport hbdiaStorageReport<float>(const HBDIA<float>&);
template HBDIAStorageReport hbdiaStorageReport<int>(const HBDIA<int>&);
