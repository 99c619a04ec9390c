// ORACLE — test infrastructure only.  Not part of the product and never linked by it.
//
// Minimal roaring-style set containers used by the CPU oracle: one container holds a subset of a
// 2^16 block as a sorted u16 array (card <= 4096), a 1024 x u64 bitset, or a run list, with the
// conversion rules of CRoaring 4.5.0 (`conanfile.py:22`, not vendored in the reference): arrays up
// to DEFAULT_MAX_SIZE = 4096 elements (`roaring_util/roaring_container.cpp:10-13`), bitsets above,
// and `runOptimize` switching to runs when they serialize smaller.  Representation never changes a
// set result; it only keeps the oracle's CPU timing representative of the reference.
//
// A `Bitmap` is a 32-bit set split into (high 16 bits = key, container) pairs, the layout of
// `roaring::Roaring` and of `CopyOnWriteBitmap` (`query_engine/copy_on_write_bitmap.h:28-147`):
// containers are shared, immutable and copied only when an operation produces a new one.
#pragma once

#include <algorithm>
#include <bit>
#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

namespace oracle {

constexpr uint32_t kArrayMax = 4096;
constexpr uint32_t kWords = 1024;

struct Container {
   enum Type : uint8_t { ARRAY = 1, BITSET = 2, RUN = 3 };
   Type type = ARRAY;
   std::vector<uint16_t> arr;   // ARRAY: sorted values
   std::vector<uint64_t> bits;  // BITSET: 1024 words
   std::vector<uint16_t> runs;  // RUN: (start, length - 1) pairs
   uint32_t card = 0;

   static Container fromSorted(const uint16_t* values, size_t n) {
      Container c;
      if (n <= kArrayMax) {
         c.type = ARRAY;
         c.arr.assign(values, values + n);
      } else {
         c.type = BITSET;
         c.bits.assign(kWords, 0);
         for (size_t i = 0; i < n; ++i) c.bits[values[i] >> 6] |= 1ULL << (values[i] & 63);
      }
      c.card = static_cast<uint32_t>(n);
      return c;
   }

   static Container fromWords(const uint64_t* words) {
      uint32_t card = 0;
      for (uint32_t w = 0; w < kWords; ++w) card += std::popcount(words[w]);
      Container c;
      c.card = card;
      if (card <= kArrayMax) {
         c.type = ARRAY;
         c.arr.reserve(card);
         for (uint32_t w = 0; w < kWords; ++w) {
            uint64_t x = words[w];
            while (x) {
               c.arr.push_back(static_cast<uint16_t>(w * 64 + std::countr_zero(x)));
               x &= x - 1;
            }
         }
      } else {
         c.type = BITSET;
         c.bits.assign(words, words + kWords);
      }
      return c;
   }

   void toWords(uint64_t* out) const {
      std::fill(out, out + kWords, 0ULL);
      switch (type) {
         case ARRAY:
            for (uint16_t v : arr) out[v >> 6] |= 1ULL << (v & 63);
            break;
         case BITSET:
            std::copy(bits.begin(), bits.end(), out);
            break;
         case RUN:
            for (size_t i = 0; i < runs.size(); i += 2) {
               const uint32_t s = runs[i];
               const uint32_t e = s + runs[i + 1] + 1;  // exclusive
               setRange(out, s, e);
            }
            break;
      }
   }

   static void setRange(uint64_t* out, uint32_t s, uint32_t e) {
      while (s < e) {
         const uint32_t w = s >> 6;
         const uint32_t lo = s & 63;
         const uint32_t hi = std::min<uint32_t>(64, lo + (e - s));
         const uint64_t mask = (hi - lo == 64) ? ~0ULL : (((1ULL << (hi - lo)) - 1) << lo);
         out[w] |= mask;
         s += hi - lo;
      }
   }

   bool contains(uint16_t v) const {
      switch (type) {
         case ARRAY:
            return std::binary_search(arr.begin(), arr.end(), v);
         case BITSET:
            return (bits[v >> 6] >> (v & 63)) & 1;
         case RUN: {
            for (size_t i = 0; i < runs.size(); i += 2) {
               if (v < runs[i]) return false;
               if (v <= static_cast<uint32_t>(runs[i]) + runs[i + 1]) return true;
            }
            return false;
         }
      }
      return false;
   }

   template <typename F>
   void forEach(F&& f) const {
      switch (type) {
         case ARRAY:
            for (uint16_t v : arr) f(v);
            break;
         case BITSET:
            for (uint32_t w = 0; w < kWords; ++w) {
               uint64_t x = bits[w];
               while (x) {
                  f(static_cast<uint16_t>(w * 64 + std::countr_zero(x)));
                  x &= x - 1;
               }
            }
            break;
         case RUN:
            for (size_t i = 0; i < runs.size(); i += 2) {
               const uint32_t s = runs[i];
               const uint32_t e = s + runs[i + 1];
               for (uint32_t v = s; v <= e; ++v) f(static_cast<uint16_t>(v));
            }
            break;
      }
   }

   /// Array or bitset equivalent of a run container (operations are defined on those two forms).
   Container normalized() const {
      if (type != RUN) return *this;
      if (card <= kArrayMax) {
         Container c;
         c.type = ARRAY;
         c.card = card;
         c.arr.reserve(card);
         forEach([&](uint16_t v) { c.arr.push_back(v); });
         return c;
      }
      Container c;
      c.type = BITSET;
      c.card = card;
      c.bits.assign(kWords, 0);
      toWords(c.bits.data());
      return c;
   }

   uint32_t countRuns() const {
      uint32_t n = 0;
      int64_t prev = -2;
      forEach([&](uint16_t v) {
         if (static_cast<int64_t>(v) != prev + 1) ++n;
         prev = v;
      });
      return n;
   }

   size_t sizeInBytes() const {
      switch (type) {
         case ARRAY:
            return 2 + 2 * arr.size();
         case BITSET:
            return 8192;
         case RUN:
            return 2 + 2 * runs.size();
      }
      return 0;
   }

   /// CRoaring run_optimize: keep the smallest of the current form and the run form.
   void runOptimize() {
      if (type == RUN || card == 0) return;
      const uint32_t n_runs = countRuns();
      const size_t run_bytes = 2 + 4 * static_cast<size_t>(n_runs);
      if (run_bytes >= sizeInBytes()) return;
      std::vector<uint16_t> r;
      r.reserve(2 * n_runs);
      int64_t start = -1, prev = -2;
      forEach([&](uint16_t v) {
         if (static_cast<int64_t>(v) != prev + 1) {
            if (start >= 0) {
               r.push_back(static_cast<uint16_t>(start));
               r.push_back(static_cast<uint16_t>(prev - start));
            }
            start = v;
         }
         prev = v;
      });
      if (start >= 0) {
         r.push_back(static_cast<uint16_t>(start));
         r.push_back(static_cast<uint16_t>(prev - start));
      }
      arr.clear();
      arr.shrink_to_fit();
      bits.clear();
      bits.shrink_to_fit();
      runs = std::move(r);
      type = RUN;
   }
};

using ContainerPtr = std::shared_ptr<const Container>;

inline uint32_t rangeCardinality(const Container& c, uint32_t s, uint32_t e_incl) {
   // number of elements of c in [s, e_incl]
   switch (c.type) {
      case Container::ARRAY: {
         auto lo = std::lower_bound(c.arr.begin(), c.arr.end(), static_cast<uint16_t>(s));
         auto hi = std::upper_bound(c.arr.begin(), c.arr.end(), static_cast<uint16_t>(e_incl));
         return static_cast<uint32_t>(hi - lo);
      }
      case Container::BITSET: {
         uint32_t n = 0;
         const uint32_t ws = s >> 6, we = e_incl >> 6;
         for (uint32_t w = ws; w <= we; ++w) {
            uint64_t x = c.bits[w];
            if (w == ws) x &= ~0ULL << (s & 63);
            if (w == we && (e_incl & 63) != 63) x &= (1ULL << ((e_incl & 63) + 1)) - 1;
            n += std::popcount(x);
         }
         return n;
      }
      case Container::RUN: {
         uint32_t n = 0;
         for (size_t i = 0; i < c.runs.size(); i += 2) {
            const uint32_t rs = c.runs[i], re = rs + c.runs[i + 1];
            const uint32_t a = std::max(rs, s), b = std::min(re, e_incl);
            if (a <= b) n += b - a + 1;
         }
         return n;
      }
   }
   return 0;
}

/// |a & b| without materializing (CRoaring container_and_cardinality).
inline uint32_t andCardinality(const Container& a, const Container& b) {
   if (a.type == Container::RUN) {
      uint32_t n = 0;
      for (size_t i = 0; i < a.runs.size(); i += 2) {
         n += rangeCardinality(b, a.runs[i], static_cast<uint32_t>(a.runs[i]) + a.runs[i + 1]);
      }
      return n;
   }
   if (b.type == Container::RUN) return andCardinality(b, a);
   if (a.type == Container::ARRAY && b.type == Container::ARRAY) {
      uint32_t n = 0;
      size_t i = 0, j = 0;
      while (i < a.arr.size() && j < b.arr.size()) {
         if (a.arr[i] < b.arr[j]) {
            ++i;
         } else if (a.arr[i] > b.arr[j]) {
            ++j;
         } else {
            ++n, ++i, ++j;
         }
      }
      return n;
   }
   if (a.type == Container::ARRAY) {
      uint32_t n = 0;
      for (uint16_t v : a.arr) n += (b.bits[v >> 6] >> (v & 63)) & 1;
      return n;
   }
   if (b.type == Container::ARRAY) return andCardinality(b, a);
   uint32_t n = 0;
   for (uint32_t w = 0; w < kWords; ++w) n += std::popcount(a.bits[w] & b.bits[w]);
   return n;
}

inline Container andC(const Container& a0, const Container& b0) {
   const Container a = a0.normalized();
   const Container b = b0.normalized();
   if (a.type == Container::ARRAY && b.type == Container::ARRAY) {
      std::vector<uint16_t> out;
      std::set_intersection(a.arr.begin(), a.arr.end(), b.arr.begin(), b.arr.end(), std::back_inserter(out));
      return Container::fromSorted(out.data(), out.size());
   }
   if (a.type == Container::ARRAY || b.type == Container::ARRAY) {
      const Container& arr = a.type == Container::ARRAY ? a : b;
      const Container& bs = a.type == Container::ARRAY ? b : a;
      std::vector<uint16_t> out;
      out.reserve(arr.arr.size());
      for (uint16_t v : arr.arr) {
         if ((bs.bits[v >> 6] >> (v & 63)) & 1) out.push_back(v);
      }
      return Container::fromSorted(out.data(), out.size());
   }
   uint64_t w[kWords];
   for (uint32_t i = 0; i < kWords; ++i) w[i] = a.bits[i] & b.bits[i];
   return Container::fromWords(w);
}

inline Container orC(const Container& a0, const Container& b0) {
   const Container a = a0.normalized();
   const Container b = b0.normalized();
   if (a.type == Container::ARRAY && b.type == Container::ARRAY) {
      std::vector<uint16_t> out;
      std::set_union(a.arr.begin(), a.arr.end(), b.arr.begin(), b.arr.end(), std::back_inserter(out));
      return Container::fromSorted(out.data(), out.size());
   }
   uint64_t wa[kWords], wb[kWords];
   a.toWords(wa);
   b.toWords(wb);
   for (uint32_t i = 0; i < kWords; ++i) wa[i] |= wb[i];
   return Container::fromWords(wa);
}

inline Container andnotC(const Container& a0, const Container& b0) {
   const Container a = a0.normalized();
   const Container b = b0.normalized();
   if (a.type == Container::ARRAY) {
      std::vector<uint16_t> out;
      out.reserve(a.arr.size());
      if (b.type == Container::ARRAY) {
         std::set_difference(a.arr.begin(), a.arr.end(), b.arr.begin(), b.arr.end(), std::back_inserter(out));
      } else {
         for (uint16_t v : a.arr) {
            if (!((b.bits[v >> 6] >> (v & 63)) & 1)) out.push_back(v);
         }
      }
      return Container::fromSorted(out.data(), out.size());
   }
   uint64_t wa[kWords], wb[kWords];
   a.toWords(wa);
   b.toWords(wb);
   for (uint32_t i = 0; i < kWords; ++i) wa[i] &= ~wb[i];
   return Container::fromWords(wa);
}

/// Complement of the bits in [lo, hi) of `c`, keeping the rest (roaring flip on one container).
inline Container flipC(const Container& c, uint32_t lo, uint32_t hi) {
   uint64_t w[kWords];
   c.toWords(w);
   uint64_t m[kWords] = {};
   Container::setRange(m, lo, hi);
   for (uint32_t i = 0; i < kWords; ++i) w[i] ^= m[i];
   return Container::fromWords(w);
}

inline Container rangeC(uint32_t lo, uint32_t hi) {
   uint64_t w[kWords] = {};
   Container::setRange(w, lo, hi);
   Container c = Container::fromWords(w);
   c.runOptimize();
   return c;
}

struct Bitmap {
   std::vector<uint16_t> keys;
   std::vector<ContainerPtr> cs;

   bool empty() const { return keys.empty(); }

   uint64_t cardinality() const {
      uint64_t n = 0;
      for (const auto& c : cs) n += c->card;
      return n;
   }

   void push(uint16_t key, Container&& c) {
      if (c.card == 0) return;
      keys.push_back(key);
      cs.push_back(std::make_shared<const Container>(std::move(c)));
   }

   void pushShared(uint16_t key, ContainerPtr c) {
      if (!c || c->card == 0) return;
      keys.push_back(key);
      cs.push_back(std::move(c));
   }

   const Container* find(uint16_t key) const {
      auto it = std::lower_bound(keys.begin(), keys.end(), key);
      if (it == keys.end() || *it != key) return nullptr;
      return cs[it - keys.begin()].get();
   }

   bool contains(uint32_t v) const {
      const Container* c = find(static_cast<uint16_t>(v >> 16));
      return c != nullptr && c->contains(static_cast<uint16_t>(v & 0xFFFF));
   }

   static Bitmap fromSorted(const uint32_t* v, size_t n) {
      Bitmap b;
      size_t i = 0;
      std::vector<uint16_t> low;
      while (i < n) {
         const uint16_t key = static_cast<uint16_t>(v[i] >> 16);
         low.clear();
         while (i < n && (v[i] >> 16) == key) {
            if (low.empty() || low.back() != static_cast<uint16_t>(v[i] & 0xFFFF)) {
               low.push_back(static_cast<uint16_t>(v[i] & 0xFFFF));
            }
            ++i;
         }
         b.push(key, Container::fromSorted(low.data(), low.size()));
      }
      return b;
   }

   std::vector<uint32_t> toVector() const {
      std::vector<uint32_t> out;
      out.reserve(cardinality());
      for (size_t i = 0; i < keys.size(); ++i) {
         const uint32_t hi = static_cast<uint32_t>(keys[i]) << 16;
         cs[i]->forEach([&](uint16_t v) { out.push_back(hi | v); });
      }
      return out;
   }

   template <typename F>
   void forEach(F&& f) const {
      for (size_t i = 0; i < keys.size(); ++i) {
         const uint32_t hi = static_cast<uint32_t>(keys[i]) << 16;
         cs[i]->forEach([&](uint16_t v) { f(hi | v); });
      }
   }

   void runOptimize() {
      for (auto& c : cs) {
         Container copy = *c;
         copy.runOptimize();
         c = std::make_shared<const Container>(std::move(copy));
      }
   }
};

inline Bitmap andB(const Bitmap& a, const Bitmap& b) {
   Bitmap r;
   size_t i = 0, j = 0;
   while (i < a.keys.size() && j < b.keys.size()) {
      if (a.keys[i] < b.keys[j]) {
         ++i;
      } else if (a.keys[i] > b.keys[j]) {
         ++j;
      } else {
         r.push(a.keys[i], andC(*a.cs[i], *b.cs[j]));
         ++i, ++j;
      }
   }
   return r;
}

inline uint64_t andCardinalityB(const Bitmap& a, const Bitmap& b) {
   uint64_t n = 0;
   size_t i = 0, j = 0;
   while (i < a.keys.size() && j < b.keys.size()) {
      if (a.keys[i] < b.keys[j]) {
         ++i;
      } else if (a.keys[i] > b.keys[j]) {
         ++j;
      } else {
         n += andCardinality(*a.cs[i], *b.cs[j]);
         ++i, ++j;
      }
   }
   return n;
}

inline Bitmap andnotB(const Bitmap& a, const Bitmap& b) {
   Bitmap r;
   size_t j = 0;
   for (size_t i = 0; i < a.keys.size(); ++i) {
      while (j < b.keys.size() && b.keys[j] < a.keys[i]) ++j;
      if (j < b.keys.size() && b.keys[j] == a.keys[i]) {
         r.push(a.keys[i], andnotC(*a.cs[i], *b.cs[j]));
      } else {
         r.pushShared(a.keys[i], a.cs[i]);
      }
   }
   return r;
}

inline Bitmap orB(const Bitmap& a, const Bitmap& b) {
   Bitmap r;
   size_t i = 0, j = 0;
   while (i < a.keys.size() || j < b.keys.size()) {
      if (j == b.keys.size() || (i < a.keys.size() && a.keys[i] < b.keys[j])) {
         r.pushShared(a.keys[i], a.cs[i]);
         ++i;
      } else if (i == a.keys.size() || b.keys[j] < a.keys[i]) {
         r.pushShared(b.keys[j], b.cs[j]);
         ++j;
      } else {
         r.push(a.keys[i], orC(*a.cs[i], *b.cs[j]));
         ++i, ++j;
      }
   }
   return r;
}

/// Flips [lo, hi) (64-bit bounds, may span several containers).
inline Bitmap flipB(const Bitmap& a, uint64_t lo, uint64_t hi) {
   if (lo >= hi) return a;
   Bitmap r;
   const uint32_t klo = static_cast<uint32_t>(lo >> 16);
   const uint32_t khi = static_cast<uint32_t>((hi - 1) >> 16);
   size_t i = 0;
   for (; i < a.keys.size() && a.keys[i] < klo; ++i) r.pushShared(a.keys[i], a.cs[i]);
   for (uint32_t k = klo; k <= khi; ++k) {
      const uint32_t s = k == klo ? static_cast<uint32_t>(lo & 0xFFFF) : 0;
      const uint32_t e = k == khi ? static_cast<uint32_t>(((hi - 1) & 0xFFFF) + 1) : 0x10000;
      if (i < a.keys.size() && a.keys[i] == k) {
         r.push(static_cast<uint16_t>(k), flipC(*a.cs[i], s, e));
         ++i;
      } else {
         r.push(static_cast<uint16_t>(k), rangeC(s, e));
      }
   }
   for (; i < a.keys.size(); ++i) r.pushShared(a.keys[i], a.cs[i]);
   return r;
}

}  // namespace oracle
