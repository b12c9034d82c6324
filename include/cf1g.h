/*
 * cf1g — MI355X (gfx950) phase API for the cf1-rs P1..P4 hot path.
 *
 * A drop-in for the phase calls of cf1-rs run_pipeline (src/pipeline.rs:66-118):
 * each entry point replaces one reference function, cited beside it. Plain C
 * types only; no exceptions or panics cross this boundary: every call returns
 * CF1G_OK or a negative error code, and cf1g_last_error(ctx) describes it.
 *
 * One context owns one HIP stream and all device buffers (sequence, bins,
 * keys, MPHF, states). It is not re-entrant; phases must run in order
 * (load -> P1 -> P2 -> P3 -> P4) and a call out of order returns
 * CF1G_E_ORDER. Every call returns after its GPU work has completed.
 *
 * k-mer keys cross the boundary as u64 for k <= 32 and as (lo, hi) u64 pairs
 * (the little-endian u128 layout of the reference's Kmer storage,
 * src/kmer.rs:407-460) for 33 <= k <= 63.
 */
#ifndef CF1G_H
#define CF1G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CF1G_OK 0
#define CF1G_E_INVALID (-1)   /* bad argument (e.g. even k: src/params.rs:167-170) */
#define CF1G_E_ORDER (-2)     /* phase called before its inputs exist */
#define CF1G_E_HIP (-3)       /* HIP runtime error, including device OOM */
#define CF1G_E_NODEVICE (-4)  /* no usable gfx950 device */
#define CF1G_E_INTERNAL (-5)  /* invariant violated (reported, never silent) */

#define CF1G_NUM_BUCKETS 65536 /* NUM_HISTOGRAM_BUCKETS, src/minimizer.rs:8 */
#define CF1G_MAX_BINS 4096

typedef struct cf1g_ctx cf1g_ctx;

typedef struct cf1g_params {
  uint32_t k;              /* odd, 1..63 (src/params.rs:167-170) */
  uint32_t num_bins;       /* --num-bins (src/params.rs:58-60); 1..4096 */
  uint32_t m;              /* minimizer length; 0 = min(k/2,15) made odd (src/params.rs:221-228) */
  uint32_t p4_chunk;       /* P4 work-stealing chunk (k-mer positions); 0 = tuned */
  uint32_t p4_wgs_per_cu;  /* P4 persistent work-groups per CU; 0 = tuned */
  uint32_t p3_target_seg;  /* P3 target k-mers per sort segment; 0 = tuned */
  uint32_t stream_chunk_log2; /* cf1g_p1_stream: log2 of the stream words (64 positions) per copied chunk; 0 = 18 */
  uint32_t reserved;
} cf1g_params;

typedef struct cf1g_info {
  uint64_t n_records;      /* records loaded */
  uint64_t stream_len;     /* flat stream length (bases + separators) */
  uint64_t n_kmers;        /* W: valid k-mer occurrences */
  uint64_t n_superkmers;   /* S: super k-mers routed */
  uint64_t n_distinct;     /* D: distinct canonical k-mers */
  uint64_t n_state_words;  /* ceil(D/12) */
  uint64_t mphf_words;     /* size of cf1g_mphf_export in u64 words */
  uint32_t num_bins;       /* bins after partitioning (<= requested) */
  uint32_t seg_bits;       /* P3 segment bits per bin */
  uint32_t mphf_big_segments; /* MPHF segments built through global memory */
  uint32_t n_oversized;    /* P3 segments sorted through global memory */
  uint32_t k, m;
  uint64_t graph_captures; /* phase launch sequences captured into hipGraphs */
  uint64_t graph_replays;  /* ... and replayed */
} cf1g_info;

/* Create / destroy a context (device 0 of the current process). */
int cf1g_create(const cf1g_params* params, cf1g_ctx** out);
void cf1g_destroy(cf1g_ctx* ctx);
const char* cf1g_last_error(const cf1g_ctx* ctx);
int cf1g_get_info(cf1g_ctx* ctx, cf1g_info* out);

/* Input records (the FASTA records P1/P2/P4 re-read: src/minimizer.rs:44-89,
 * src/superkmer.rs:44-91, src/classify.rs:540-597). ascii[rec_off[r] ..
 * rec_off[r+1]) is record r; any byte outside ACGTacgt is a placeholder
 * (src/dna.rs:81-92). Host buffers are copied to HBM. */
int cf1g_load_records(cf1g_ctx* ctx, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec);

/* Host ingest — the FASTA reading that the reference repeats in every phase
 * (needletail::parse_fastx_file at src/minimizer.rs:44-50,
 * src/superkmer.rs:44-50, src/classify.rs:540-546, src/output.rs:166-172),
 * done once: the files (plain FASTA, or gzip when the name ends in .gz) are
 * parsed in order by `threads` host threads (0 = 1), the 2-bit stream is
 * packed on the host and only the packed stream (a quarter of the bases'
 * bytes plus one bit per base) is copied to HBM. A record is a '>' header
 * (id = the text up to the first whitespace, src/classify.rs:552-557) and
 * its sequence lines without line breaks. The records stay in the context
 * for P5 (cf1g_loaded_records). An unreadable file returns CF1G_E_INVALID.
 * out may be NULL. */
typedef struct cf1g_ingest_stats {
  uint64_t n_records, bases;
  uint64_t stream_bytes; /* packed stream copied to HBM */
  double parse_ms, pack_ms, h2d_ms;
} cf1g_ingest_stats;
int cf1g_load_fasta(cf1g_ctx* ctx, const char* const* paths, uint32_t n_paths, uint32_t threads, cf1g_ingest_stats* out);
/* The records of the last cf1g_load_fasta, owned by the context (valid until
 * the next load or cf1g_destroy): the arguments cf1g_p5_write takes. Any
 * output may be NULL. */
int cf1g_loaded_records(cf1g_ctx* ctx, const uint8_t** ascii, const uint64_t** rec_off, uint64_t* n_rec, const char** names,
                        const uint64_t** name_off);
/* The packed stream in HBM (the bases P1..P5 read): ceil(stream_len / 64)
 * words of placeholder bits (MSB-first, one bit per position) and twice as
 * many words of 2-bit bases (32 per word, MSB-first; placeholders 0).
 * Available after cf1g_load_fasta, or after P1 for cf1g_load_records. Either
 * output may be NULL. */
int cf1g_stream_export(cf1g_ctx* ctx, uint64_t* seq, uint64_t* nmask, uint64_t n_words);

/* P1 — count_minimizer_histogram (src/minimizer.rs:27-98): one count per super
 * k-mer at bucket_hash(minimizer) (src/minimizer.rs:18-21). out_hist may be
 * NULL; else it receives CF1G_NUM_BUCKETS counts. */
int cf1g_p1_histogram(cf1g_ctx* ctx, uint64_t* out_hist);

/* P1 straight from host records — cf1g_load_records + cf1g_p1_histogram in
 * one call, with the records' transfer overlapped: `threads` host threads
 * (0 = all) pack chunks of the 2-bit stream into pinned staging buffers, each
 * chunk's copy to HBM overlaps the packing of the next, and the P1 scan runs
 * on the tiles whose bases have landed. The same records (ascii, rec_off as
 * for cf1g_load_records, read only during the call) give the same histogram
 * and the same later products. What the reference's P1 reads from the FASTA
 * records (src/minimizer.rs:27-98). out_hist may be NULL. */
int cf1g_p1_stream(cf1g_ctx* ctx, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, uint32_t threads, uint64_t* out_hist);

/* P2 — partition_minimizers (src/minimizer.rs:116-144) + route_superkmers
 * (src/superkmer.rs:21-101). hist == NULL uses P1's device histogram;
 * otherwise the caller's CF1G_NUM_BUCKETS counts are partitioned. Optional
 * outputs: the bucket->bin table (CF1G_NUM_BUCKETS u16, Partitioning,
 * src/minimizer.rs:262-293) and the resulting bin count. */
int cf1g_p2_partition_route(cf1g_ctx* ctx, const uint64_t* hist, uint32_t num_bins, uint16_t* out_bucket_to_bin,
                            uint32_t* out_num_bins);

/* P3 — Mphf::<K>::build (src/mphf.rs:32-43): per-bin dedup (dedup_bin,
 * src/mphf.rs:323-363) then the BBHash over all D keys; also allocates the
 * zeroed state vector (AtomicStateVector::new, src/state_vector.rs:17-24).
 * out_D (may be NULL) = Mphf::total_kmers (src/mphf.rs:278-280). */
int cf1g_p3_build(cf1g_ctx* ctx, uint64_t* out_D);

/* P4 — classify_vertices::<K> (src/classify.rs:524-601). P3 joined each
 * distinct k-mer's per-occurrence observations; this encodes the joined
 * states into the state vector. */
int cf1g_p4_classify(cf1g_ctx* ctx);

/* P4 in the reference's form — classify_vertices::<K> (src/classify.rs:524-601)
 * with Classifier::process_* (src/classify.rs:25-519) and
 * AtomicStateVector::compare_and_swap (src/state_vector.rs:47-70): the state
 * vector is zeroed, then every k-mer occurrence of the loaded records is
 * looked up in the MPHF (Mphf::hash, src/mphf.rs:268-270) and joined into its
 * 5-bit state by a 64-bit compare-and-swap loop. A persistent launch
 * work-steals fixed-size chunks of positions (cf1g_params.p4_chunk,
 * p4_wgs_per_cu; 0 = the tuned values). The states equal cf1g_p4_classify's
 * word for word. flags: CF1G_CAS_FORCE_RETRY fails each occurrence's first
 * compare-and-swap once (the retry path, as the reference's
 * compare_exchange_weak may fail spuriously). out may be NULL. */
#define CF1G_CAS_FORCE_RETRY 1u
typedef struct cf1g_cas_stats {
  uint64_t occurrences;  /* k-mer occurrences classified (= cf1g_info.n_kmers) */
  uint64_t cas_issued;   /* 64-bit compare-and-swap operations issued */
  uint64_t cas_failed;   /* of which failed (a concurrent update or a forced retry) */
  uint64_t n_chunks;     /* work-stealing chunks */
  uint32_t chunk;        /* stream positions per chunk */
  uint32_t wgs_per_cu;   /* persistent work-groups per CU */
  uint32_t grid;         /* work-groups launched */
  uint32_t reserved;
  double kernel_ms;      /* device time of the classification kernel */
} cf1g_cas_stats;
int cf1g_p4_classify_cas(cf1g_ctx* ctx, uint32_t flags, cf1g_cas_stats* out);

/* Autotuner (no reference counterpart: the reference takes --num-bins from
 * the command line, src/params.rs:58-60, and sizes its tasks by fixed
 * constants, src/classify.rs:576-594). On the loaded records it times
 * bin counts {128, 256, 512} x P3 segment targets {1024, 1536, 2048} over
 * whole P1..P4 steps, and P4 chunks {2048, 8192, 32768} x work-groups per CU
 * {4, 8, 16} over the CAS classification, `reps` timed runs each after a
 * warm-up (minimum kept). The context then uses the fastest values (they
 * replace num_bins and p3_target_seg of its params) and holds the products
 * of the chosen configuration. entries (may be NULL) receives up to cap
 * measurements. Bin count and segment target change the MPHF's partition
 * (exported), never the graph: keys, states per k-mer and P5 output. */
#define CF1G_TUNE_STEP 0u   /* a = bins, b = P3 segment target, c = bins after partitioning */
#define CF1G_TUNE_P4_CAS 1u /* a = chunk, b = work-groups per CU */
typedef struct cf1g_tune_entry {
  uint32_t kind, a, b, c;
  double ms;
} cf1g_tune_entry;
typedef struct cf1g_tune_result {
  uint32_t num_bins, p3_target_seg, p4_chunk, p4_wgs_per_cu;
  double step_ms;   /* best P1..P4 step */
  double p4_cas_ms; /* best CAS classification kernel */
  uint32_t n_entries, reserved;
} cf1g_tune_result;
int cf1g_autotune(cf1g_ctx* ctx, uint32_t reps, cf1g_tune_entry* entries, uint32_t cap, cf1g_tune_result* out);

/* P1..P4 in order on the loaded records; out_D may be NULL. */
int cf1g_run_p1_p4(cf1g_ctx* ctx, uint64_t* out_D);
/* P1..P4 from host records: cf1g_p1_stream, then P2..P4 (run_pipeline's
 * phases from records in host memory to the state vector,
 * src/pipeline.rs:77-103). */
int cf1g_run_p1_p4_stream(cf1g_ctx* ctx, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, uint32_t threads,
                          uint64_t* out_D);

/* ---- products consumed by the host P5 */

/* AtomicStateVector words (12 x 5-bit states per u64, src/state_vector.rs:7-42). */
int cf1g_states_export(cf1g_ctx* ctx, uint64_t* words, uint64_t n_words);
/* The distinct canonical keys (the key set of cf1_dedup_all.bin,
 * src/mphf.rs:120-132) in MPHF value order: key v has value v. D u64, or D
 * (lo, hi) pairs. */
int cf1g_keys_export(cf1g_ctx* ctx, uint64_t* out, uint64_t n_words);
/* Serialised MPHF (minimizer-partitioned BBHash): 16 header words
 * (magic, k, m, segment bits, bins, segments, D, blocks, wide), the
 * bucket->bin table (65536 u16), per-segment key offsets and block offsets
 * (segments + 1 words each), then 8-word blocks {rank, 7 bit words}. With
 * buf == NULL only *n_words is set. The layout is complete: a lookup over
 * it with the segment and level hashes the device uses gives exactly
 * cf1g_mphf_lookup's values (Mphf::hash, src/mphf.rs:268-270). */
int cf1g_mphf_export(cf1g_ctx* ctx, uint64_t* buf, uint64_t* n_words);
/* Batch Mphf::hash (src/mphf.rs:268-270): keys in host memory (u64 or pairs). */
int cf1g_mphf_lookup(cf1g_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out);
/* No counterpart in the reference: cf1g_mphf_lookup for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same keys in host memory, same values,
 * but the keys and values are staged through a fixed 32 MiB device array
 * that each device holds for the life of the process, so the call allocates
 * nothing; at most 32 MiB / (key bytes + 8) - 64 keys per call (2,097,088 for
 * k <= 32, 1,398,037 for k > 32), more returns CF1G_E_INVALID. Calls on one
 * device share that array: calls that overlap on one device (from several
 * threads or contexts) race. */
int cf1g_mphf_lookup_noalloc(cf1g_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out);
/* Same on device-resident keys and output (n keys). */
int cf1g_mphf_lookup_device(cf1g_ctx* ctx, const void* d_keys, uint64_t n, uint64_t* d_out);
/* Per stream position (cf1g_info.stream_len bytes): bit0 valid k-mer,
 * bit1 forward (DirectedKmer::dir, src/directed_kmer.rs:26-36), bits 2-3 state
 * class (StateClass, src/state.rs:6-11) — what AnnotatedKmer::new computes per
 * k-mer for P5 (src/directed_kmer.rs:81-96). */
int cf1g_annotate(cf1g_ctx* ctx, uint8_t* out, uint64_t n);
/* Maximal unitigs, marked on the GPU (is_unipath_start / is_unipath_end of
 * src/output.rs:79-145 over every position, the unitig id of
 * output_gfa_unitig :714-736 through the selected index, and its outputted
 * claim :747-751 by compare-and-swap on a copy of the state vector). In
 * stream order: start / end = stream positions of the first and last k-mer
 * (record-local index = position - cf1g_record_starts()[r]); id = the
 * unitig's id; flags bit0 = forward (first k-mer < reverse complement of the
 * last), bit1 = this unitig claimed its id (the one whose line cf_seg holds).
 * *n_spans receives the count; with all outputs NULL nothing else is
 * written; otherwise cap must be >= the count. The state vector is not
 * changed. */
int cf1g_unitig_spans(cf1g_ctx* ctx, uint64_t* n_spans, uint64_t* start, uint64_t* end, uint64_t* id, uint8_t* flags, uint64_t cap);
/* The input file of each loaded record (0-based), for cf_seq's
 * "Reference:<file+1>" (ref_id = file_idx + 1 over params.input_files,
 * src/output.rs:176-177). cf1g_load_fasta sets it from its paths;
 * cf1g_load_records sets every record to file 0, and a caller that loaded
 * records of several files gives them here (n_rec entries). */
int cf1g_set_record_files(cf1g_ctx* ctx, const uint32_t* file_idx, uint64_t n_rec);
int cf1g_record_files(cf1g_ctx* ctx, uint32_t* out, uint64_t n);
/* Stream position of record r's first base (n_rec + 1 entries; the last is the
 * stream length). */
int cf1g_record_starts(cf1g_ctx* ctx, uint64_t* out, uint64_t n);
/* Routed bin records (P2 product, fixed-width form of the bin record format
 * of src/superkmer.rs:181-202): words per record (2 or 4) via *rec_words,
 * bin boundaries (num_bins + 1 record offsets) and the records. Any output
 * may be NULL. */
int cf1g_bins_export(cf1g_ctx* ctx, uint32_t* rec_words, uint64_t* bin_start, uint64_t* records, uint64_t n_words);

/* ---- the reference's intermediate files, and resume */
/* Routed super k-mers as the reference's bin files (P2 product): dir/cf1_bin_%05d.bin
 * per bin (bin_file_path, src/superkmer.rs:224-226), records of a u16 LE base
 * count and ceil(n/4) bytes with base j at bits 6-2(j%4) of byte j/4
 * (pack_and_buffer, src/superkmer.rs:181-210; BinReader :259-289). Records
 * of a bin are in device order (the reference's order depends on its thread
 * schedule). threads (0 = 1) write bins in parallel; out_bytes may be NULL. */
int cf1g_bins_write(cf1g_ctx* ctx, const char* dir, uint32_t threads, uint64_t* out_bytes);
/* The consolidated distinct keys as the reference writes them in P3
 * (cf1_dedup_all.bin, src/mphf.rs:120-132 from dedup_bin :323-363): each
 * bin's canonical keys sorted, bins in order, raw native-endian storage (u64,
 * or u128 with the low word first for k > 32). */
int cf1g_dedup_write(cf1g_ctx* ctx, const char* path, uint32_t threads);
/* Resume (no reference counterpart; the reference's phase_bench reuses P1..P3
 * intermediates ad hoc): cf1g_save writes the MPHF (the cf1g_mphf_export
 * words) and the state words after P4; cf1g_restore, on a context with the
 * same records loaded, reads them back and leaves the context after P4:
 * lookups, annotation, cf1g_p4_classify_cas and P5 work, while products that
 * need the keys or the joined observations (keys, dedup file, learned index,
 * cf1g_p4_classify) return CF1G_E_ORDER until P1..P3 run again. A file of
 * other records, k or layout returns CF1G_E_INVALID. */
int cf1g_save(cf1g_ctx* ctx, const char* path);
int cf1g_restore(cf1g_ctx* ctx, const char* path);

/* ---- host P5 (stays on the host, per the reference's phase split)
 * extract_and_output + write_json (src/output.rs:152-354, 785-834): writes
 * prefix.cf_seg / .cf_seq / .json for the records given to cf1g_load_records
 * (same ascii/rec_off), using the GPU's per-position annotation, states and
 * MPHF. names_blob[name_off[r] .. name_off[r+1]) is record r's id (may be
 * NULL). out_meta5 (may be NULL): unitigs, vertices, max, min, sum length. */
int cf1g_p5_write(cf1g_ctx* ctx, const uint8_t* ascii, const uint64_t* rec_off, uint64_t n_rec, const char* names_blob,
                  const uint64_t* name_off, const char* out_prefix, const char* input_desc, int poly_n_stretch,
                  int track_short_seqs, int threads, uint64_t* out_meta5);

/* ---- learned rank index (no reference counterpart: an alternative to
 * Mphf::hash, src/mphf.rs:268-276, whose values are the ranks of the keys in
 * the globally sorted distinct k-mers — a monotone minimal perfect map).
 * u64 keys only (k <= 31). Trained on the GPU after P3. */
typedef struct cf1g_lidx_params {
  uint32_t leaf_bits;  /* log2 L; 0 = 17 (capped at floor(log2 D)); L + R leaves in all */
  uint32_t root_bits;  /* log2 root pieces R; 0 = 8 (at most 20; up to 8 the lookup holds them in LDS) */
  uint32_t epochs;     /* gradient-descent epochs per level; 0 = 4 */
  float lr;            /* step size on standardised inputs; 0 = 0.5 */
  uint32_t max_window; /* leaves with a wider error window resolve through the BBHash; 0 = 4096 */
  uint32_t untrained;  /* 1: keep the endpoint interpolation (the method without training) */
  uint32_t line_bytes; /* 64 or 128; 0 = 64 */
  uint32_t line_keys;  /* keys per line C; 0 = 256 (128-byte lines) or 108 (64-byte lines) */
} cf1g_lidx_params;

typedef struct cf1g_lidx_info {
  uint64_t n_keys, n_leaves, n_pieces;
  uint64_t fallback_leaves, fallback_keys; /* leaves (and their keys) resolved through the BBHash */
  uint64_t n_lines, overflow_lines;        /* lines; lines too full for their codes (searched) */
  uint64_t search_keys;                    /* keys in runs the codes leave to a search of the keys */
  uint32_t line_bytes, line_keys;
  uint32_t max_err_lo, max_err_hi;         /* recorded error bounds (non-fallback leaves) */
  double mean_abs_err;                     /* mean |rank - prediction| over all keys */
  double line_bits_per_key;                /* bits the lines' codes use, per key (overflow lines left out) */
  uint64_t bytes_total;                    /* root + leaf records + lines */
  double root_loss[2], leaf_loss[2];       /* MSE before the first and after the last epoch */
  double sort_ms, train_ms;                /* sort of the keys; whole build (sort included) */
} cf1g_lidx_info;

/* Sort the distinct keys, train root and leaves (forward / backward / update
 * kernels), record error bounds, build the lines. Needs P3. */
int cf1g_lidx_train(cf1g_ctx* ctx, const cf1g_lidx_params* params, cf1g_lidx_info* out);
/* Exact rank of each present key; an absent key gets 0 or some rank, as with
 * any minimal perfect map (src/mphf.rs:268-270). Host buffers /
 * device-resident buffers. */
int cf1g_lidx_lookup(cf1g_ctx* ctx, const uint64_t* keys, uint64_t n, uint64_t* out);
int cf1g_lidx_lookup_device(cf1g_ctx* ctx, const void* d_keys, uint64_t n, uint64_t* d_out);
/* The trained index: its shape, the sorted keys (D), the key warp's tables
 * (warp_groups x 256 u32 cumulative byte frequencies x 2^-32), root pieces
 * (R + 1 records of 16 bytes: f32 A, f32 B, u32 first leaf, u32 first rank;
 * piece = warped key >> root_shift), leaf records (n_leaves x 8 bytes: f32 slope whose sign
 * bit marks a BBHash leaf, f32 intercept), the leaves' error bounds
 * (n_leaves x {u32 below, u32 above}) and the lines (n_lines x line_bytes).
 * Any output may be NULL. */
typedef struct cf1g_lidx_shape {
  uint64_t n_keys;
  uint32_t n_pieces, n_leaves;
  int32_t root_shift;
  uint32_t line_bytes, line_keys, n_lines;
  uint32_t key_bits, warp_groups;
} cf1g_lidx_shape;
int cf1g_lidx_export(cf1g_ctx* ctx, cf1g_lidx_shape* shape, uint64_t* sorted_keys, uint32_t* warp, void* root, void* leaves,
                     uint32_t* leaf_err, void* lines);
/* Gradient check: with the caller's parameters (w, c in standardised
 * coordinates, one pair per segment; level 0 = root pieces, 1 = leaves) run
 * the forward and backward kernels and return per-segment gradients and
 * summed squared residuals, plus the segment starts (n_segs + 1) and the
 * standardisation (mu, 1/deviation) the kernels used. Outputs may be NULL. */
int cf1g_lidx_gradients(cf1g_ctx* ctx, int level, const float* w, const float* c, uint64_t n_segs, double* out_gw, double* out_gc,
                        double* out_sq, uint64_t* out_seg, float* out_mu, float* out_isd);
/* P4 in the reference's form through an index (0 = BBHash, 1 = learned):
 * one lookup per k-mer occurrence of the records, in sequence order, and a
 * CAS join of its observation into a state vector in the index's order
 * (classify_vertices, src/classify.rs:524-601, with Mphf::hash replaced by
 * the index). The BBHash form is cf1g_p4_classify_cas's; out (may be NULL)
 * as there. */
int cf1g_classify_via(cf1g_ctx* ctx, int index, cf1g_cas_stats* out);
/* Which index cf1g_annotate, cf1g_states_export and cf1g_p5_write use
 * (0 = BBHash, the default; 1 = learned, after cf1g_classify_via(ctx, 1)). */
int cf1g_select_index(cf1g_ctx* ctx, int index);
int cf1g_selected_index(const cf1g_ctx* ctx);

/* ---- execution and measurement */
/* Each phase's kernel launches between host synchronisations run as one
 * hipGraph, captured on first use and replayed while the sizes and device
 * buffers it was captured with are unchanged (on by default; 0 launches
 * them one by one). No reference counterpart (the reference runs rayon
 * tasks, src/pipeline.rs:66-118). */
int cf1g_set_graphs(cf1g_ctx* ctx, int on);
int cf1g_sync(cf1g_ctx* ctx);
/* Enable per-kernel HIP-event timing on the context stream (0/1); phases
 * then launch kernel by kernel (no graphs). */
int cf1g_set_profiling(cf1g_ctx* ctx, int on);
/* Kernel timing summary as text lines "name count total_ms" ('\n'-separated);
 * returns the needed buffer size (including the NUL). Resets when reset != 0. */
size_t cf1g_kernel_times(cf1g_ctx* ctx, char* buf, size_t cap, int reset);
/* Device-resident copies for benchmarks that time from HBM-resident input:
 * the device pointers of the loaded ASCII buffer and of the keys array. */
const void* cf1g_device_keys(cf1g_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* CF1G_H */
