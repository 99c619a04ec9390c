// Seeded synthetic sequence/metadata generator for the SILO query-engine workloads.
//
// Produces rows already in the reference-diffed form that
// `extractCoverageAndMutationsFromSequence` (reference: src/rhydb/common/aligned_sequence.cpp:18-122)
// yields for an aligned sequence: a covered range [start, end), the internal missing (N) runs inside
// it, and the mutations (position, symbol != reference symbol, symbol != N).  The same generator can
// also re-materialize any row as the aligned string it was diffed from, so the diff itself can be
// checked against an independent restatement of the reference's diff on small inputs.
//
// Models (all distributions explicit, every row drawn from its own RNG stream so generation is
// parallel and any single row can be regenerated):
//   * plain model: Poisson(sub_lambda) substitutions at uniform positions, each to a uniform base
//     different from the reference base;
//   * lineage model: a random recursive tree of n_nodes lineages (parent of node i uniform in [0,i)),
//     Zipf(zipf_lineage) node sizes over a seeded rank permutation, Poisson(node_lambda) substitutions
//     per node over its parent, Poisson(private_lambda) private substitutions per row;
//   * both: with probability del_prob one deletion run of '-' of length U{1..del_max};
//     Poisson(n_lambda) N runs with 1 + Geometric(1/n_mean) lengths (mean n_mean), applied last;
//   * metadata: country (Zipf(zipf_country) over n_countries values), date (uniform over n_days),
//     lineage id (the row's tree node).
// Edits apply in the order substitutions, deletion, N runs (later ones overwrite earlier ones).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <utility>
#include <vector>

namespace {

constexpr char kNucChars[] = "-ACGTRYSWKMBDHVN";
constexpr uint8_t kGap = 0;
constexpr uint8_t kMissing = 15;

uint64_t splitmix64(uint64_t x) {
   x += 0x9E3779B97F4A7C15ULL;
   x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
   x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
   return x ^ (x >> 31);
}

struct Params {
   uint32_t n_rows;
   uint32_t genome_length;
   uint64_t seed;
   int lineage_model;
   double sub_lambda;
   uint32_t n_nodes;
   double zipf_lineage;
   double node_lambda;
   double private_lambda;
   double del_prob;
   uint32_t del_max;
   double n_lambda;
   double n_mean;
   uint32_t n_countries;
   double zipf_country;
   uint32_t n_days;
   int32_t date_base;
   double null_seq_frac;
   double null_meta_frac;
};

struct Edits {
   bool is_null = false;
   std::vector<std::pair<uint32_t, uint8_t>> subs;  // applied in order
   uint32_t del_start = 0, del_end = 0;              // [start, end)
   std::vector<std::pair<uint32_t, uint32_t>> n_runs;  // [start, end), unmerged
   uint32_t country = 0, lineage = 0;
   int32_t date = 0;
   bool country_null = false, date_null = false, lineage_null = false;
};

struct Dataset {
   Params p;
   std::vector<uint8_t> reference;
   // lineage tree
   std::vector<int32_t> parent;
   std::vector<std::vector<std::pair<uint32_t, uint8_t>>> node_subs;  // sorted by position
   std::vector<double> lineage_weight;
   std::vector<double> country_weight;
   std::vector<uint32_t> country_name_perm, lineage_name_perm;
   // diffed rows
   std::vector<uint8_t> is_null;
   std::vector<uint32_t> cov_start, cov_end;
   std::vector<uint64_t> mut_ptr;
   std::vector<uint32_t> mut_pos;
   std::vector<uint8_t> mut_sym;
   std::vector<uint64_t> nrun_ptr;
   std::vector<uint32_t> nrun_start, nrun_end;
   // metadata
   std::vector<uint32_t> country, lineage;
   std::vector<int32_t> date;
   std::vector<uint8_t> country_null, date_null, lineage_null;
   std::string names_buf;
   // amino-acid columns (dg_translate)
   struct AaColumn {
      std::vector<uint8_t> reference;
      std::vector<uint32_t> cov_start, cov_end;
      std::vector<uint64_t> mut_ptr, nrun_ptr;
      std::vector<uint32_t> mut_pos, nrun_start, nrun_end;
      std::vector<uint8_t> mut_sym;
   };
   std::vector<AaColumn> aa;
};

uint8_t otherBase(std::mt19937_64& rng, uint8_t current) {
   // uniform over the three bases A,C,G,T (1..4) different from `current`
   std::uniform_int_distribution<int> pick(0, 2);
   int k = pick(rng);
   uint8_t base = 1;
   for (uint8_t b = 1; b <= 4; ++b) {
      if (b == current) continue;
      if (k-- == 0) {
         base = b;
         break;
      }
   }
   return base;
}

std::vector<double> zipfWeights(uint32_t n, double s) {
   std::vector<double> w(n);
   for (uint32_t k = 0; k < n; ++k) w[k] = 1.0 / std::pow(static_cast<double>(k + 1), s);
   return w;
}

Edits drawEdits(const Dataset& ds, uint32_t row) {
   const Params& p = ds.p;
   std::mt19937_64 rng(splitmix64(p.seed ^ splitmix64(0xA11CE5ULL + row)));
   Edits e;
   std::uniform_real_distribution<double> unit(0.0, 1.0);
   std::uniform_int_distribution<uint32_t> pos_dist(0, p.genome_length - 1);
   e.is_null = p.null_seq_frac > 0 && unit(rng) < p.null_seq_frac;
   if (p.lineage_model) {
      std::discrete_distribution<uint32_t> lin(ds.lineage_weight.begin(), ds.lineage_weight.end());
      e.lineage = lin(rng);
      const auto& base_subs = ds.node_subs[e.lineage];
      e.subs = base_subs;
      std::poisson_distribution<int> npriv(p.private_lambda);
      const int k = npriv(rng);
      for (int i = 0; i < k; ++i) {
         const uint32_t pos = pos_dist(rng);
         uint8_t current = ds.reference[pos];
         auto it = std::lower_bound(
            base_subs.begin(), base_subs.end(), std::make_pair(pos, static_cast<uint8_t>(0))
         );
         if (it != base_subs.end() && it->first == pos) current = it->second;
         e.subs.emplace_back(pos, otherBase(rng, current));
      }
   } else {
      std::poisson_distribution<int> nsub(p.sub_lambda);
      const int k = nsub(rng);
      for (int i = 0; i < k; ++i) {
         const uint32_t pos = pos_dist(rng);
         e.subs.emplace_back(pos, otherBase(rng, ds.reference[pos]));
      }
   }
   if (unit(rng) < p.del_prob) {
      std::uniform_int_distribution<uint32_t> len(1, p.del_max);
      const uint32_t start = pos_dist(rng);
      e.del_start = start;
      e.del_end = std::min<uint32_t>(p.genome_length, start + len(rng));
   }
   std::poisson_distribution<int> nn(p.n_lambda);
   std::geometric_distribution<uint32_t> geo(1.0 / p.n_mean);
   const int kn = nn(rng);
   for (int i = 0; i < kn; ++i) {
      const uint32_t start = pos_dist(rng);
      const uint64_t end = static_cast<uint64_t>(start) + 1 + geo(rng);
      e.n_runs.emplace_back(start, static_cast<uint32_t>(std::min<uint64_t>(end, p.genome_length)));
   }
   std::discrete_distribution<uint32_t> ctry(ds.country_weight.begin(), ds.country_weight.end());
   e.country = ctry(rng);
   std::uniform_int_distribution<int32_t> day(0, static_cast<int32_t>(p.n_days) - 1);
   e.date = p.date_base + day(rng);
   if (p.null_meta_frac > 0) {
      e.country_null = unit(rng) < p.null_meta_frac;
      e.date_null = unit(rng) < p.null_meta_frac;
      e.lineage_null = unit(rng) < p.null_meta_frac;
   }
   return e;
}

struct RowDiff {
   uint32_t start = 0, end = 0;
   std::vector<std::pair<uint32_t, uint32_t>> internal_n;
   std::vector<std::pair<uint32_t, uint8_t>> mutations;
};

std::vector<std::pair<uint32_t, uint32_t>> mergeRuns(std::vector<std::pair<uint32_t, uint32_t>> runs) {
   std::sort(runs.begin(), runs.end());
   std::vector<std::pair<uint32_t, uint32_t>> merged;
   for (const auto& r : runs) {
      if (!merged.empty() && r.first <= merged.back().second) {
         merged.back().second = std::max(merged.back().second, r.second);
      } else {
         merged.push_back(r);
      }
   }
   return merged;
}

RowDiff diffEdits(const Dataset& ds, const Edits& e) {
   RowDiff d;
   if (e.is_null) return d;
   const uint32_t L = ds.p.genome_length;
   auto n = mergeRuns(e.n_runs);
   uint32_t start = 0, end = L;
   size_t first = 0, last = n.size();
   if (!n.empty() && n.front().first == 0) {
      start = n.front().second;
      first = 1;
   }
   if (last > first && n.back().second == L) {
      end = n.back().first;
      last -= 1;
   }
   if (start >= end) {  // the whole sequence is missing
      return d;
   }
   d.start = start;
   d.end = end;
   d.internal_n.assign(n.begin() + static_cast<ptrdiff_t>(first), n.begin() + static_cast<ptrdiff_t>(last));
   auto in_n = [&](uint32_t pos) {
      auto it = std::upper_bound(
         d.internal_n.begin(), d.internal_n.end(), std::make_pair(pos, UINT32_MAX)
      );
      return it != d.internal_n.begin() && pos < std::prev(it)->second;
   };
   // final symbol at every edited position: last substitution, then deletion on top
   std::vector<std::pair<uint32_t, uint8_t>> edits;
   edits.reserve(e.subs.size() + (e.del_end - e.del_start));
   for (size_t i = 0; i < e.subs.size(); ++i) edits.push_back(e.subs[i]);
   std::stable_sort(edits.begin(), edits.end(), [](const auto& a, const auto& b) {
      return a.first < b.first;
   });
   std::vector<std::pair<uint32_t, uint8_t>> last_sub;
   for (const auto& ed : edits) {
      if (!last_sub.empty() && last_sub.back().first == ed.first) {
         last_sub.back().second = ed.second;
      } else {
         last_sub.push_back(ed);
      }
   }
   std::vector<std::pair<uint32_t, uint8_t>> final_sym;
   size_t si = 0;
   for (uint32_t pos = e.del_start; pos < e.del_end; ++pos) {
      while (si < last_sub.size() && last_sub[si].first < pos) final_sym.push_back(last_sub[si++]);
      if (si < last_sub.size() && last_sub[si].first == pos) ++si;
      final_sym.emplace_back(pos, kGap);
   }
   while (si < last_sub.size()) final_sym.push_back(last_sub[si++]);
   for (const auto& [pos, sym] : final_sym) {
      if (pos < start || pos >= end || in_n(pos)) continue;
      if (sym == ds.reference[pos]) continue;
      d.mutations.emplace_back(pos, sym);
   }
   return d;
}

void buildTree(Dataset& ds, std::mt19937_64& rng) {
   const Params& p = ds.p;
   const uint32_t n = p.n_nodes;
   ds.parent.assign(n, -1);
   ds.node_subs.assign(n, {});
   std::uniform_int_distribution<uint32_t> pos_dist(0, p.genome_length - 1);
   std::poisson_distribution<int> nsub(p.node_lambda);
   for (uint32_t i = 0; i < n; ++i) {
      std::vector<std::pair<uint32_t, uint8_t>> subs;
      if (i > 0) {
         std::uniform_int_distribution<uint32_t> par(0, i - 1);
         ds.parent[i] = static_cast<int32_t>(par(rng));
         subs = ds.node_subs[ds.parent[i]];
      }
      const int k = nsub(rng);
      for (int j = 0; j < k; ++j) {
         const uint32_t pos = pos_dist(rng);
         auto it = std::lower_bound(subs.begin(), subs.end(), std::make_pair(pos, static_cast<uint8_t>(0)));
         const bool present = it != subs.end() && it->first == pos;
         const uint8_t current = present ? it->second : ds.reference[pos];
         const uint8_t base = otherBase(rng, current);
         if (present) {
            it->second = base;
         } else {
            subs.insert(it, {pos, base});
         }
      }
      ds.node_subs[i] = std::move(subs);
   }
   auto w = zipfWeights(n, p.zipf_lineage);
   std::vector<uint32_t> rank(n);
   for (uint32_t i = 0; i < n; ++i) rank[i] = i;
   std::shuffle(rank.begin(), rank.end(), rng);
   ds.lineage_weight.assign(n, 0.0);
   for (uint32_t i = 0; i < n; ++i) ds.lineage_weight[i] = w[rank[i]];
}

}  // namespace

extern "C" {

/// Generates a dataset.  Returns an opaque handle (free with dg_free) or nullptr on bad parameters.
void* dg_generate(
   uint32_t n_rows,
   uint32_t genome_length,
   uint64_t seed,
   int lineage_model,
   double sub_lambda,
   uint32_t n_nodes,
   double zipf_lineage,
   double node_lambda,
   double private_lambda,
   double del_prob,
   uint32_t del_max,
   double n_lambda,
   double n_mean,
   uint32_t n_countries,
   double zipf_country,
   uint32_t n_days,
   int32_t date_base,
   double null_seq_frac,
   double null_meta_frac
) {
   if (genome_length == 0 || n_countries == 0 || n_days == 0 || n_mean < 1.0 ||
       (lineage_model && n_nodes == 0) || del_max == 0) {
      return nullptr;
   }
   auto* ds = new Dataset;
   ds->p = Params{n_rows,      genome_length, seed,        lineage_model, sub_lambda,   n_nodes,
                  zipf_lineage, node_lambda,  private_lambda, del_prob,   del_max,      n_lambda,
                  n_mean,      n_countries,   zipf_country, n_days,     date_base,    null_seq_frac,
                  null_meta_frac};
   std::mt19937_64 rng(splitmix64(seed));
   std::uniform_int_distribution<int> base(1, 4);
   ds->reference.resize(genome_length);
   for (auto& b : ds->reference) b = static_cast<uint8_t>(base(rng));
   if (lineage_model) {
      buildTree(*ds, rng);
   } else {
      ds->parent.assign(1, -1);
      ds->node_subs.assign(1, {});
      ds->lineage_weight.assign(1, 1.0);
   }
   ds->country_weight = zipfWeights(n_countries, zipf_country);
   // names: a seeded permutation so that string order differs from dictionary-id order
   auto perm = [&](uint32_t n) {
      std::vector<uint32_t> v(n);
      for (uint32_t i = 0; i < n; ++i) v[i] = i;
      std::shuffle(v.begin(), v.end(), rng);
      return v;
   };
   ds->country_name_perm = perm(n_countries);
   ds->lineage_name_perm = perm(static_cast<uint32_t>(ds->parent.size()));

   std::vector<RowDiff> diffs(n_rows);
   std::vector<Edits> meta(n_rows);
#pragma omp parallel for schedule(dynamic, 4096)
   for (int64_t r = 0; r < static_cast<int64_t>(n_rows); ++r) {
      Edits e = drawEdits(*ds, static_cast<uint32_t>(r));
      diffs[r] = diffEdits(*ds, e);
      e.subs.clear();
      e.subs.shrink_to_fit();
      e.n_runs.clear();
      e.n_runs.shrink_to_fit();
      meta[r] = std::move(e);
   }
   ds->is_null.resize(n_rows);
   ds->cov_start.resize(n_rows);
   ds->cov_end.resize(n_rows);
   ds->mut_ptr.assign(n_rows + 1, 0);
   ds->nrun_ptr.assign(n_rows + 1, 0);
   ds->country.resize(n_rows);
   ds->date.resize(n_rows);
   ds->lineage.resize(n_rows);
   ds->country_null.resize(n_rows);
   ds->date_null.resize(n_rows);
   ds->lineage_null.resize(n_rows);
   for (uint32_t r = 0; r < n_rows; ++r) {
      ds->mut_ptr[r + 1] = ds->mut_ptr[r] + diffs[r].mutations.size();
      ds->nrun_ptr[r + 1] = ds->nrun_ptr[r] + diffs[r].internal_n.size();
   }
   ds->mut_pos.resize(ds->mut_ptr[n_rows]);
   ds->mut_sym.resize(ds->mut_ptr[n_rows]);
   ds->nrun_start.resize(ds->nrun_ptr[n_rows]);
   ds->nrun_end.resize(ds->nrun_ptr[n_rows]);
#pragma omp parallel for schedule(static)
   for (int64_t r = 0; r < static_cast<int64_t>(n_rows); ++r) {
      const auto& d = diffs[r];
      const auto& m = meta[r];
      ds->is_null[r] = m.is_null ? 1 : 0;
      ds->cov_start[r] = d.start;
      ds->cov_end[r] = d.end;
      uint64_t o = ds->mut_ptr[r];
      for (const auto& [pos, sym] : d.mutations) {
         ds->mut_pos[o] = pos;
         ds->mut_sym[o] = sym;
         ++o;
      }
      o = ds->nrun_ptr[r];
      for (const auto& [s, e] : d.internal_n) {
         ds->nrun_start[o] = s;
         ds->nrun_end[o] = e;
         ++o;
      }
      ds->country[r] = m.country;
      ds->date[r] = m.date;
      ds->lineage[r] = m.lineage;
      ds->country_null[r] = m.country_null;
      ds->date_null[r] = m.date_null;
      ds->lineage_null[r] = m.lineage_null;
   }
   return ds;
}

void dg_free(void* handle) { delete static_cast<Dataset*>(handle); }

/// Amino-acid columns derived from the nucleotide rows: gene g covers the codons at nucleotide
/// positions gene_start[g] + 3 i .. + 2 for residues i < gene_len[g].  A residue is missing ('X')
/// when any of its three nucleotides is outside the row's covered range or inside a missing run;
/// a deletion ('-') when all three are '-'; also missing when only some are; otherwise the codon's
/// amino acid in the standard genetic code, a mutation when it differs from the reference's.
/// Leading and trailing missing residues bound the residue coverage, the others become missing
/// runs, as for nucleotides.  Null rows stay null.  Outputs are in the generator's CSR layout per
/// gene and stay owned by the dataset (dg_aa_* getters).
int dg_translate(void* handle, uint32_t n_genes, const uint32_t* gene_start, const uint32_t* gene_len);
uint32_t dg_aa_genes(void* h);
const uint8_t* dg_aa_reference(void* h, uint32_t g);
const uint32_t* dg_aa_cov_start(void* h, uint32_t g);
const uint32_t* dg_aa_cov_end(void* h, uint32_t g);
const uint64_t* dg_aa_mut_ptr(void* h, uint32_t g);
const uint32_t* dg_aa_mut_pos(void* h, uint32_t g);
const uint8_t* dg_aa_mut_sym(void* h, uint32_t g);
uint64_t dg_aa_num_mutations(void* h, uint32_t g);
const uint64_t* dg_aa_nrun_ptr(void* h, uint32_t g);
const uint32_t* dg_aa_nrun_start(void* h, uint32_t g);
const uint32_t* dg_aa_nrun_end(void* h, uint32_t g);
uint64_t dg_aa_num_nruns(void* h, uint32_t g);

uint32_t dg_num_rows(void* h) { return static_cast<Dataset*>(h)->p.n_rows; }
uint32_t dg_genome_length(void* h) { return static_cast<Dataset*>(h)->p.genome_length; }
uint32_t dg_num_lineages(void* h) { return static_cast<uint32_t>(static_cast<Dataset*>(h)->parent.size()); }
uint64_t dg_num_mutations(void* h) { return static_cast<Dataset*>(h)->mut_pos.size(); }
uint64_t dg_num_nruns(void* h) { return static_cast<Dataset*>(h)->nrun_start.size(); }

const uint8_t* dg_reference(void* h) { return static_cast<Dataset*>(h)->reference.data(); }
const uint8_t* dg_is_null(void* h) { return static_cast<Dataset*>(h)->is_null.data(); }
const uint32_t* dg_cov_start(void* h) { return static_cast<Dataset*>(h)->cov_start.data(); }
const uint32_t* dg_cov_end(void* h) { return static_cast<Dataset*>(h)->cov_end.data(); }
const uint64_t* dg_mut_ptr(void* h) { return static_cast<Dataset*>(h)->mut_ptr.data(); }
const uint32_t* dg_mut_pos(void* h) { return static_cast<Dataset*>(h)->mut_pos.data(); }
const uint8_t* dg_mut_sym(void* h) { return static_cast<Dataset*>(h)->mut_sym.data(); }
const uint64_t* dg_nrun_ptr(void* h) { return static_cast<Dataset*>(h)->nrun_ptr.data(); }
const uint32_t* dg_nrun_start(void* h) { return static_cast<Dataset*>(h)->nrun_start.data(); }
const uint32_t* dg_nrun_end(void* h) { return static_cast<Dataset*>(h)->nrun_end.data(); }
const uint32_t* dg_country(void* h) { return static_cast<Dataset*>(h)->country.data(); }
const int32_t* dg_date(void* h) { return static_cast<Dataset*>(h)->date.data(); }
const uint32_t* dg_lineage(void* h) { return static_cast<Dataset*>(h)->lineage.data(); }
const uint8_t* dg_country_null(void* h) { return static_cast<Dataset*>(h)->country_null.data(); }
const uint8_t* dg_date_null(void* h) { return static_cast<Dataset*>(h)->date_null.data(); }
const uint8_t* dg_lineage_null(void* h) { return static_cast<Dataset*>(h)->lineage_null.data(); }
const int32_t* dg_lineage_parent(void* h) { return static_cast<Dataset*>(h)->parent.data(); }
const double* dg_lineage_weight(void* h) { return static_cast<Dataset*>(h)->lineage_weight.data(); }
const double* dg_country_weight(void* h) { return static_cast<Dataset*>(h)->country_weight.data(); }
const uint32_t* dg_country_name_perm(void* h) { return static_cast<Dataset*>(h)->country_name_perm.data(); }
const uint32_t* dg_lineage_name_perm(void* h) { return static_cast<Dataset*>(h)->lineage_name_perm.data(); }

/// Writes the aligned sequence string of row `row` (genome_length characters, no terminator) as it
/// was before diffing: substitutions, then the deletion run, then the N runs.  A null row is
/// written as all 'N'.  Returns 0 on success.
int dg_materialize(void* h, uint32_t row, char* out) {
   auto* ds = static_cast<Dataset*>(h);
   if (row >= ds->p.n_rows) return -1;
   const uint32_t L = ds->p.genome_length;
   const Edits e = drawEdits(*ds, row);
   for (uint32_t i = 0; i < L; ++i) out[i] = kNucChars[ds->reference[i]];
   if (e.is_null) {
      std::memset(out, 'N', L);
      return 0;
   }
   for (const auto& [pos, sym] : e.subs) out[pos] = kNucChars[sym];
   for (uint32_t pos = e.del_start; pos < e.del_end; ++pos) out[pos] = kNucChars[kGap];
   for (const auto& [s, t] : e.n_runs) {
      for (uint32_t pos = s; pos < t; ++pos) out[pos] = kNucChars[kMissing];
   }
   return 0;
}

/// Renders rows [row0, row0 + n) of a column in diff form as aligned strings of L characters each
/// (out: n * L): the reference's characters, the mutations, and `missing_char` outside the covered
/// range and inside the missing runs (a null row: all missing).  `chars` maps symbols to
/// characters; pointers are the column's whole arrays (mut_ptr and nrun_ptr index them globally).
void dg_render(uint32_t L, const uint8_t* reference, const char* chars, char missing_char, uint64_t row0, uint64_t n,
               const uint8_t* is_null, const uint32_t* cov_start, const uint32_t* cov_end, const uint64_t* mut_ptr,
               const uint32_t* mut_pos, const uint8_t* mut_sym, const uint64_t* nrun_ptr, const uint32_t* nrun_start,
               const uint32_t* nrun_end, char* out) {
   std::string ref(L, ' ');
   for (uint32_t i = 0; i < L; ++i) ref[i] = chars[reference[i]];
#pragma omp parallel for schedule(dynamic, 256)
   for (int64_t k = 0; k < static_cast<int64_t>(n); ++k) {
      const uint64_t r = row0 + static_cast<uint64_t>(k);
      char* o = out + static_cast<uint64_t>(k) * L;
      if (is_null[r]) {
         std::memset(o, missing_char, L);
         continue;
      }
      const uint32_t s = std::min(cov_start[r], L), e = std::min(std::max(cov_end[r], s), L);
      std::memset(o, missing_char, s);
      std::memcpy(o + s, ref.data() + s, e - s);
      std::memset(o + e, missing_char, L - e);
      for (uint64_t m = mut_ptr[r]; m < mut_ptr[r + 1]; ++m) o[mut_pos[m]] = chars[mut_sym[m]];
      for (uint64_t m = nrun_ptr[r]; m < nrun_ptr[r + 1]; ++m) {
         for (uint32_t p = nrun_start[m]; p < nrun_end[m] && p < L; ++p) o[p] = missing_char;
      }
   }
}

}  // extern "C"

namespace {

// Standard genetic code; codon index 16 b1 + 4 b2 + b3 with T = 0, C = 1, A = 2, G = 3.
constexpr char kCode[] = "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
constexpr char kAaChars[] = "-ACDEFGHIKLMNOPQRSTUVWYBJZ*X";
constexpr uint8_t kAaGap = 0;
constexpr uint8_t kAaMissing = 27;

uint8_t aaIndex(char c) {
   for (uint8_t i = 0; kAaChars[i]; ++i) {
      if (kAaChars[i] == c) return i;
   }
   return kAaMissing;
}

// nucleotide symbol (A=1, C=2, G=3, T=4) -> code index digit (T=0, C=1, A=2, G=3)
int codeDigit(uint8_t nuc) {
   switch (nuc) {
      case 4: return 0;
      case 2: return 1;
      case 1: return 2;
      case 3: return 3;
      default: return -1;
   }
}

uint8_t translate(uint8_t a, uint8_t b, uint8_t c) {
   if (a == kGap && b == kGap && c == kGap) return kAaGap;
   const int x = codeDigit(a), y = codeDigit(b), z = codeDigit(c);
   if (x < 0 || y < 0 || z < 0) return kAaMissing;  // partial deletion or an ambiguous base
   return aaIndex(kCode[16 * x + 4 * y + z]);
}

}  // namespace

extern "C" {

int dg_translate(void* handle, uint32_t n_genes, const uint32_t* gene_start, const uint32_t* gene_len) {
   auto* ds = static_cast<Dataset*>(handle);
   const uint32_t L = ds->p.genome_length;
   const uint32_t n = ds->p.n_rows;
   for (uint32_t g = 0; g < n_genes; ++g) {
      if (gene_len[g] == 0 || gene_start[g] + 3ULL * gene_len[g] > L) return -1;
   }
   ds->aa.assign(n_genes, {});
   for (uint32_t g = 0; g < n_genes; ++g) {
      auto& col = ds->aa[g];
      const uint32_t gs = gene_start[g], gl = gene_len[g];
      col.reference.resize(gl);
      for (uint32_t i = 0; i < gl; ++i) {
         col.reference[i] = translate(ds->reference[gs + 3 * i], ds->reference[gs + 3 * i + 1], ds->reference[gs + 3 * i + 2]);
      }
      // Row r's residues in the gene, in time proportional to its mutations and missing runs.  A
      // residue is missing when its codon is not fully inside the covered range [cs, ce), touches
      // an internal N run, or translates to the missing symbol (partial deletion, ambiguous base).
      // Leading and trailing missing residues bound the amino-acid covered range; the missing
      // residues inside it form its runs; the other residues that differ from the gene's
      // reference are its mutations.  Returns false for a row without a covered residue.
      struct Scratch {
         std::vector<std::pair<uint32_t, uint32_t>> miss;  // missing residue intervals
         std::vector<std::pair<uint32_t, uint8_t>> muts;
      };
      auto rowAa = [&](uint32_t r, Scratch& w, uint32_t& start, uint32_t& end) -> bool {
         w.miss.clear();
         w.muts.clear();
         start = end = 0;
         if (ds->is_null[r] || ds->cov_start[r] >= ds->cov_end[r]) return false;
         const uint32_t cs = ds->cov_start[r], ce = ds->cov_end[r];
         // residues whose codon lies inside [cs, ce): [lo, hi)
         const uint32_t lo = cs <= gs ? 0u : std::min<uint32_t>(gl, (cs - gs + 2) / 3);
         const uint32_t hi = ce < gs ? 0u : std::min<uint32_t>(gl, (ce - gs) / 3);
         if (lo >= hi) return false;
         for (uint64_t k = ds->nrun_ptr[r]; k < ds->nrun_ptr[r + 1]; ++k) {
            const uint32_t a = ds->nrun_start[k], b = ds->nrun_end[k];
            if (b <= gs || a >= gs + 3 * gl) continue;
            const uint32_t i0 = a > gs ? (a - gs) / 3 : 0;
            const uint32_t i1 = std::min<uint32_t>(gl, (b - gs + 2) / 3);
            w.miss.emplace_back(i0, i1);
         }
         // the runs are disjoint and ascending in nucleotides, so both ends of their residue
         // intervals are non-decreasing: the last interval starting at or before i is the one
         // that can cover i
         const size_t n_run_iv = w.miss.size();
         auto inRun = [&](uint32_t i) {
            auto it = std::upper_bound(w.miss.begin(), w.miss.begin() + n_run_iv, std::make_pair(i, UINT32_MAX));
            return it != w.miss.begin() && std::prev(it)->second > i;
         };
         const uint32_t* mp = ds->mut_pos.data() + ds->mut_ptr[r];
         const uint32_t* me = ds->mut_pos.data() + ds->mut_ptr[r + 1];
         const uint8_t* ms = ds->mut_sym.data() + ds->mut_ptr[r];
         auto nucAt = [&](uint32_t pos) -> uint8_t {
            const uint32_t* it = std::lower_bound(mp, me, pos);
            return (it != me && *it == pos) ? ms[it - mp] : ds->reference[pos];
         };
         uint32_t last = UINT32_MAX;
         for (const uint32_t* it = std::lower_bound(mp, me, gs); it != me && *it < gs + 3 * gl; ++it) {
            const uint32_t i = (*it - gs) / 3;
            if (i == last) continue;
            last = i;
            if (i < lo || i >= hi || inRun(i)) continue;
            const uint32_t a = gs + 3 * i;
            const uint8_t aa = translate(nucAt(a), nucAt(a + 1), nucAt(a + 2));
            if (aa == kAaMissing) {
               w.miss.emplace_back(i, i + 1);
            } else if (aa != col.reference[i]) {
               w.muts.emplace_back(i, aa);
            }
         }
         // clip to [lo, hi), merge overlapping or adjacent intervals
         std::sort(w.miss.begin(), w.miss.end());
         size_t m = 0;
         for (size_t k = 0; k < w.miss.size(); ++k) {
            const uint32_t a = std::max(w.miss[k].first, lo), b = std::min(w.miss[k].second, hi);
            if (a >= b) continue;
            if (m > 0 && a <= w.miss[m - 1].second) {
               w.miss[m - 1].second = std::max(w.miss[m - 1].second, b);
            } else {
               w.miss[m++] = {a, b};
            }
         }
         w.miss.resize(m);
         uint32_t s0 = lo, e0 = hi;
         size_t first = 0, stop = m;
         if (m > 0 && w.miss[0].first <= s0) s0 = w.miss[first++].second;
         if (stop > first && w.miss[stop - 1].second >= e0) e0 = w.miss[--stop].first;
         if (s0 >= e0) {
            w.miss.clear();
            w.muts.clear();
            return false;
         }
         start = s0;
         end = e0;
         w.miss.erase(w.miss.begin() + stop, w.miss.end());
         w.miss.erase(w.miss.begin(), w.miss.begin() + first);
         return true;
      };
      col.cov_start.assign(n, 0);
      col.cov_end.assign(n, 0);
      col.mut_ptr.assign(n + 1, 0);
      col.nrun_ptr.assign(n + 1, 0);
#pragma omp parallel
      {
         Scratch w;
#pragma omp for schedule(dynamic, 4096)
         for (int64_t r = 0; r < static_cast<int64_t>(n); ++r) {
            uint32_t st, en;
            if (rowAa(static_cast<uint32_t>(r), w, st, en)) {
               col.cov_start[r] = st;
               col.cov_end[r] = en;
               col.mut_ptr[r + 1] = w.muts.size();
               col.nrun_ptr[r + 1] = w.miss.size();
            }
         }
      }
      for (uint32_t r = 0; r < n; ++r) {
         col.mut_ptr[r + 1] += col.mut_ptr[r];
         col.nrun_ptr[r + 1] += col.nrun_ptr[r];
      }
      col.mut_pos.resize(col.mut_ptr[n]);
      col.mut_sym.resize(col.mut_ptr[n]);
      col.nrun_start.resize(col.nrun_ptr[n]);
      col.nrun_end.resize(col.nrun_ptr[n]);
#pragma omp parallel
      {
         Scratch w;
#pragma omp for schedule(dynamic, 4096)
         for (int64_t r = 0; r < static_cast<int64_t>(n); ++r) {
            uint32_t st, en;
            if (!rowAa(static_cast<uint32_t>(r), w, st, en)) continue;
            uint64_t o = col.mut_ptr[r];
            for (const auto& [i, a] : w.muts) {
               col.mut_pos[o] = i;
               col.mut_sym[o] = a;
               ++o;
            }
            o = col.nrun_ptr[r];
            for (const auto& [a, b] : w.miss) {
               col.nrun_start[o] = a;
               col.nrun_end[o] = b;
               ++o;
            }
         }
      }
   }
   return 0;
}

uint32_t dg_aa_genes(void* h) { return static_cast<uint32_t>(static_cast<Dataset*>(h)->aa.size()); }
const uint8_t* dg_aa_reference(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).reference.data(); }
const uint32_t* dg_aa_cov_start(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).cov_start.data(); }
const uint32_t* dg_aa_cov_end(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).cov_end.data(); }
const uint64_t* dg_aa_mut_ptr(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).mut_ptr.data(); }
const uint32_t* dg_aa_mut_pos(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).mut_pos.data(); }
const uint8_t* dg_aa_mut_sym(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).mut_sym.data(); }
uint64_t dg_aa_num_mutations(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).mut_pos.size(); }
const uint64_t* dg_aa_nrun_ptr(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).nrun_ptr.data(); }
const uint32_t* dg_aa_nrun_start(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).nrun_start.data(); }
const uint32_t* dg_aa_nrun_end(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).nrun_end.data(); }
uint64_t dg_aa_num_nruns(void* h, uint32_t g) { return static_cast<Dataset*>(h)->aa.at(g).nrun_start.size(); }

}  // extern "C"
