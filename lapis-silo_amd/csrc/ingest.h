// Ingestion on the device: aligned sequences diffed by a HIP kernel into the builder's diff form
// (SequenceColumnBuilder, storage/column/sequence_column.h:178-233), and the device stages of a
// sequence column's index build (local-reference adaptation and sorted container entries).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "engine.h"
#include "kernels.h"

namespace rdb {

// The device stages of buildSequenceColumn over a column's rows in diff form, runs normalized and
// packed start | end << 16.  The rows are uploaded once, by the constructor.
class DeviceIndexBuild {
  public:
   DeviceIndexBuild(Table& t, const SeqColumn& col, const uint8_t* is_null, const uint32_t* cov_start,
                    const uint32_t* cov_end, const uint64_t* mut_ptr, const uint32_t* mut_pos, const uint8_t* mut_sym,
                    const uint64_t* nrun_ptr, const uint32_t* runs);
   // Covered rows and the adapted local reference symbol per position.
   void adapt(std::vector<uint32_t>& coverage, std::vector<uint8_t>& local_ref);
   // Every chunk's container entries, sorted: keys[chunk_ptr[c] .. chunk_ptr[c + 1]) are chunk c's,
   // each c << 40 | position << 24 | symbol << 16 | row in chunk.
   void entries(const std::vector<uint8_t>& flip_to, const std::vector<uint32_t>& flipped, RawVector<uint64_t>& keys,
                std::vector<uint64_t>& chunk_ptr);

  private:
   Table& t_;
   DiffRowsDev rows_{};
   uint64_t mut_total_ = 0, run_total_ = 0;
   std::vector<DeviceBuffer> bufs_;
};

// A sequence column under construction from aligned sequences: each append is diffed on the device
// and kept in diff form; finish builds the column into the table.
struct SeqBuilder {
   Table* table = nullptr;
   std::unique_ptr<SeqColumn> col;
   bool finished = false;
   std::vector<uint8_t> is_null;
   std::vector<uint32_t> cov_start, cov_end;
   std::vector<uint64_t> mut_ptr{0}, nrun_ptr{0};
   std::vector<uint32_t> mut_pos, nrun_start, nrun_end;
   std::vector<uint8_t> mut_sym;
   DeviceBuffer lut, ref;
};

// Diffs n_rows aligned sequences (row i: data[row_ptr[i] .. row_ptr[i + 1]) from position
// offsets[i], or 0 when offsets is null; is_null may be null) and appends them; all or nothing.
void seqBuilderAppend(SeqBuilder& b, const char* data, const uint64_t* row_ptr, const uint32_t* offsets,
                      const uint8_t* is_null, uint64_t n_rows);
// Builds the column (every row of the table appended) and adds it to the table.
void seqBuilderFinish(SeqBuilder& b);

// 64-bit digest of a sequence column's device buffers and host-side index state.
uint64_t seqColumnDigest(const Table& t, const SeqColumn& col);

}  // namespace rdb
