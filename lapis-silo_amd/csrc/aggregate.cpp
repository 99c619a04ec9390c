// Grouping dimensions and histogram decoding of the Aggregated action
// (operators/bitmap_aggregation_node.cpp:53-356), shared by the single-query and batched paths.
#include <algorithm>
#include <cstdlib>

#include "capi_internal.h"

namespace rdb::capi {

std::vector<AggDim> parseDims(const Table& t, const char* dims) {
   std::vector<AggDim> ds;
   const std::string spec(dims ? dims : "");
   size_t b = 0;
   while (b <= spec.size()) {
      size_t e = spec.find(',', b);
      if (e == std::string::npos) e = spec.size();
      const std::string d = spec.substr(b, e - b);
      b = e + 1;
      if (d.empty()) continue;
      if (ds.size() == 4) fail(RDB_EQUERY, "at most four grouping dimensions are supported");
      const auto at = d.find('@');
      AggDim dim;
      if (at != std::string::npos) {
         const int idx = t.seqIndex(d.substr(0, at));
         if (idx < 0) fail(RDB_EQUERY, "Database does not contain the sequence '" + d.substr(0, at) + "'");
         const SeqColumn& col = *t.seq[idx];
         char* endp = nullptr;
         const unsigned long p = std::strtoul(d.c_str() + at + 1, &endp, 10);
         if (*endp != '\0' || p >= col.genome_length) fail(RDB_EQUERY, "bad sequence position in '" + d + "'");
         dim.is_seq = true;
         dim.seq = &col;
         dim.seq_index = idx;
         dim.position = static_cast<uint32_t>(p);
         // bitmap_aggregation_node.cpp:53-84: every symbol, plus the null group
         dim.n_labels = col.abc->count + 1;
      } else {
         const DictColumn* col = t.dictColumn(d);
         if (!col) fail(RDB_EQUERY, "The database does not contain the column '" + d + "'");
         dim.dict = col;
         dim.n_labels = static_cast<uint32_t>(col->dictionary.size()) + 1;
      }
      ds.push_back(dim);
   }
   return ds;
}

uint64_t dimStrides(const std::vector<AggDim>& ds, uint32_t* stride) {
   uint64_t bins = 1;
   for (int k = static_cast<int>(ds.size()) - 1; k >= 0; --k) {
      stride[k] = static_cast<uint32_t>(bins);
      bins *= ds[k].n_labels;
      if (bins > (1ULL << 28)) fail(RDB_EQUERY, "too many group combinations");
   }
   return bins;
}

void decodeHistogram(const std::vector<AggDim>& ds, const uint32_t* stride, const uint32_t* hist, uint64_t bins,
                     std::vector<uint32_t>& ids, std::vector<uint64_t>& counts) {
   std::vector<std::pair<uint32_t, uint32_t>> nonzero;
   for (uint64_t bin = 0; bin < bins; ++bin) {
      if (hist[bin]) nonzero.emplace_back(static_cast<uint32_t>(bin), hist[bin]);
   }
   decodeBins(ds, stride, nonzero, ids, counts);
}

void decodeBins(const std::vector<AggDim>& ds, const uint32_t* stride, const std::vector<std::pair<uint32_t, uint32_t>>& bins,
                std::vector<uint32_t>& ids, std::vector<uint64_t>& counts) {
   struct Combo {
      uint32_t key[4];
      uint32_t order[4];
      uint64_t count;
   };
   std::vector<Combo> combos;
   for (const auto& [bin, count] : bins) {
      if (!count) continue;
      Combo c{};
      uint64_t rem = bin;
      for (size_t k = 0; k < ds.size(); ++k) {
         const uint32_t lab = static_cast<uint32_t>(rem / stride[k]);
         rem %= stride[k];
         const bool is_null = lab == ds[k].n_labels - 1;
         c.key[k] = is_null ? kInvalidId : lab;
         c.order[k] = is_null ? kInvalidId : (ds[k].is_seq ? lab : ds[k].dict->rank[lab]);
      }
      c.count = count;
      combos.push_back(c);
   }
   std::sort(combos.begin(), combos.end(), [&](const Combo& a, const Combo& b) {
      for (size_t k = 0; k < ds.size(); ++k) {
         if (a.order[k] != b.order[k]) return a.order[k] < b.order[k];
      }
      return false;
   });
   ids.clear();
   counts.clear();
   for (const auto& c : combos) {
      for (size_t k = 0; k < ds.size(); ++k) ids.push_back(c.key[k]);
      counts.push_back(c.count);
   }
}

}  // namespace rdb::capi
