/*
 * warpstate_amd — C ABI of the MI355X (gfx950) scan engine.
 *
 * This is the drop-in boundary for warpstate's GPU scan path. The reference
 * crate exposes the path through Rust types; every entry point below names the
 * reference interface it replaces (paths relative to the reference checkout).
 *
 * Conventions
 *  - All offsets are u32 byte offsets, as in the reference (`Match` is
 *    {u32 pattern_id, u32 start, u32 end}, end exclusive;
 *    core/src/dfa/mod.rs:416-419). One scan covers at most 4 GiB - 1 MiB
 *    (WS_ERR_INPUT_TOO_LARGE beyond).
 *  - Output lists are sorted by (start, pattern_id, end)
 *    (core/src/cpu/mod.rs:67-78); batch output by
 *    (source_id, start, pattern_id, end) (core/src/batch.rs:159-165).
 *  - The caller owns inputs and outputs; the library owns device tables,
 *    workspaces and streams. Every call is blocking on its context's stream.
 *  - A context serialises calls made on it; use one context per host thread
 *    for concurrent scans (each context has its own HIP stream).
 *  - Errors: every entry point returns a ws_status; ws_last_error() returns
 *    the calling thread's last message. Results never silently change
 *    semantics: a full output buffer is WS_ERR_MATCH_BUFFER_OVERFLOW with the
 *    required count in *count (the reference falls back to a CPU rescan,
 *    core/src/gpu/dispatch.rs:243-273; this library does not).
 */
#ifndef WARPSTATE_AMD_H
#define WARPSTATE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WS_ABI_VERSION 1u

/* Error variants of core/src/error.rs:6-91, one code each. */
typedef enum ws_status {
  WS_OK = 0,
  WS_ERR_EMPTY_PATTERN_SET = 1,
  WS_ERR_EMPTY_PATTERN = 2,
  WS_ERR_NO_GPU_ADAPTER = 3,
  WS_ERR_GPU_DEVICE = 4,
  WS_ERR_BUFFER_MAP_FAILED = 5,
  WS_ERR_PATTERN_COMPILATION_FAILED = 6,
  WS_ERR_DFA_OUT_OF_BOUNDS = 7,
  WS_ERR_PATTERN_TOO_LARGE = 8,
  WS_ERR_PATTERN_SET_TOO_LARGE = 9,
  WS_ERR_INPUT_TOO_LARGE = 10,
  WS_ERR_MATCH_BUFFER_OVERFLOW = 11,
  WS_ERR_PATHOLOGICAL_REGEX = 12,
  WS_ERR_INVALID_ARGUMENT = 13
} ws_status;

typedef struct ws_ctx ws_ctx;
typedef struct ws_patterns ws_patterns;

/* matchkit::Match (core/src/dfa/mod.rs:416-419). */
typedef struct ws_match {
  uint32_t pattern_id;
  uint32_t start;
  uint32_t end;
} ws_match;

/* batch::TaggedMatch (core/src/batch.rs:22-27); offsets relative to the item. */
typedef struct ws_tagged_match {
  uint64_t source_id;
  uint32_t pattern_id;
  uint32_t start;
  uint32_t end;
  uint32_t reserved;
} ws_tagged_match;

/* batch::ScanItem (core/src/batch.rs:13-18). */
typedef struct ws_item {
  uint64_t id;
  const uint8_t* data;
  uint64_t len;
} ws_item;

enum { WS_PATTERN_LITERAL = 0, WS_PATTERN_REGEX = 1 };

/* One PatternSetBuilder::literal / ::regex call (core/src/pattern/compiler.rs:13-93). */
typedef struct ws_pattern {
  uint32_t kind;        /* WS_PATTERN_LITERAL or WS_PATTERN_REGEX */
  const uint8_t* bytes; /* literal bytes, or regex source text */
  uint64_t len;
  const char* name;     /* optional, may be NULL */
} ws_pattern;

/* PatternSetBuilder flags (core/src/pattern/compiler.rs:73-84). */
enum { WS_FLAG_CASE_INSENSITIVE = 1u };

/* Match semantics.
 * WS_MODE_OVERLAPPING: the GPU contract — every occurrence of every literal
 *   (PatternSet::scan_overlapping, core/src/cpu/scan.rs:235-247,300-349) plus
 *   leftmost-first non-overlapping regex matches per pattern.
 * WS_MODE_STANDARD: PatternSet::scan (core/src/pattern/mod.rs:112-140): the
 *   literal reduction that ScanStrategy::select picks
 *   (core/src/specialize.rs:101-179, core/src/hash_scan.rs:321-363) plus the
 *   same regex matches. */
enum { WS_MODE_OVERLAPPING = 0, WS_MODE_STANDARD = 1 };

/* Where a buffer lives. */
enum { WS_MEM_HOST = 0, WS_MEM_DEVICE = 1 };

/* AutoMatcherConfig subset (core/src/config.rs:4-61). Zero fields take defaults:
 * max_matches 1,000,000 (config.rs:35), chunk_size 128 MiB, chunk_overlap 4096. */
typedef struct ws_config {
  uint64_t max_matches;
  uint64_t chunk_size;
  uint64_t chunk_overlap;
  uint32_t flags;
  uint32_t reserved;
} ws_config;

/* Kernel choice. The autotuner and the learned router produce these. */
enum {
  WS_VARIANT_AUTO = 0,
  WS_VARIANT_HASH = 1,       /* rolling-hash prefilter, LDS bloom + L2 buckets */
  WS_VARIANT_DFA_GLOBAL = 2, /* dense factor DFA, table served from L2/HBM */
  WS_VARIANT_DFA_LDS = 3     /* dense factor DFA, table resident in LDS */
};

/* ws_tuning.flags: WS_TUNING_CHECKED runs the host-checked pipeline (the
 * host reads counts back between passes and grows buffers on demand) instead
 * of the device-sized one (every pass sized on the device, one
 * synchronisation per call; overflowing buffers fall back to the checked
 * pipeline automatically). Results are identical. */
enum { WS_TUNING_CHECKED = 1u };

typedef struct ws_tuning {
  uint32_t variant;       /* WS_VARIANT_* */
  uint32_t bytes_per_lane;/* 16, 32 or 64 (hash); 0 = default */
  uint32_t blocks_per_cu; /* persistent grid = CUs * blocks_per_cu; 0 = default */
  uint32_t flags;         /* WS_TUNING_* */
} ws_tuning;

/* Compiled-set facts (router features; core/src/pattern/ir.rs:6-33). */
typedef struct ws_patterns_info {
  uint32_t pattern_count;
  uint32_t literal_count;
  uint32_t regex_count;
  uint32_t factor_count;
  uint32_t min_factor_len;
  uint32_t max_factor_len;
  uint32_t min_literal_len;
  uint32_t max_pattern_len;   /* longest match any pattern can produce; 0xFFFFFFFF if unbounded */
  uint32_t dfa_states;        /* factor DFA states */
  uint32_t dfa_classes;       /* factor DFA byte classes */
  uint32_t regex_dfa_states;  /* sum over regex anchored DFAs */
  uint32_t hash_window;       /* K of the rolling-hash prefilter */
  uint32_t variants_supported;/* bit (1 << WS_VARIANT_*) */
  uint32_t reserved;
} ws_patterns_info;

/* Per-call timings measured with HIP events on the context stream (ms). */
typedef struct ws_timings {
  float prefilter_ms;  /* the streaming prefilter kernel */
  float verify_ms;     /* exact verification of prefilter hits (hash variant) */
  float regex_ms;      /* anchored walks + chain resolution */
  float order_ms;      /* sort + pack */
  float total_ms;      /* whole call, first launch to last */
  float reserved;
  uint64_t hits;       /* prefilter hits (hash variant) */
  uint64_t candidates; /* regex candidate starts */
  uint64_t raw_matches;/* matches before ordering */
} ws_timings;

/* ---- context: GpuMatcher::new / with_config (core/src/gpu/mod.rs:95-220) ---- */
ws_status ws_ctx_create(int device_ordinal, const ws_config* cfg, ws_ctx** out);
void ws_ctx_destroy(ws_ctx* ctx);
/* The context's HIP stream (hipStream_t) for callers that enqueue around it. */
void* ws_ctx_stream(ws_ctx* ctx);
/* Number of compute units of the context's device. */
uint32_t ws_ctx_cu_count(ws_ctx* ctx);
/* Captured launch sequences the context holds. An overlapping-mode scan of
 * device memory that repeats (same patterns, input, output and tuning) is
 * captured into a hipGraph on its second call and replayed from then on. */
uint32_t ws_ctx_graph_count(ws_ctx* ctx);
/* Turn that capture and replay off (every call launched eagerly; results are
 * the same) or back on (the default). For profilers and debugging. */
void ws_ctx_set_graphs(ws_ctx* ctx, int on);
/* Enable per-call HIP-event timing (ws_ctx_last_timings). */
void ws_ctx_set_profiling(ws_ctx* ctx, int on);
ws_status ws_ctx_last_timings(ws_ctx* ctx, ws_timings* out);
const char* ws_last_error(void);
uint32_t ws_abi_version(void);

/* ---- patterns: PatternSetBuilder::build (core/src/pattern/compiler.rs:97-340) ----
 * The compiled set is uploaded to the context's device and is read-only from
 * then on: any context on that device may scan with it, concurrently. */
ws_status ws_patterns_compile(ws_ctx* ctx, const ws_pattern* patterns, size_t count,
                              uint32_t flags, ws_patterns** out);
/* A compiled set is reference counted: compile returns one reference, retain adds one
 * (Arc<PatternSet>::clone), free drops one and releases the set with the last. */
void ws_patterns_retain(ws_patterns* patterns);
void ws_patterns_free(ws_patterns* patterns);
/* Host-only compile (no device needed): validates the set exactly like
 * ws_patterns_compile and reports its facts; the learned router's features. */
ws_status ws_patterns_analyze(const ws_pattern* patterns, size_t count, uint32_t flags,
                              ws_patterns_info* out);
ws_status ws_patterns_info_get(const ws_patterns* patterns, ws_patterns_info* out);

/* ---- scan: GpuMatcher::scan (core/src/gpu/mod.rs:223-297) + dispatch::scan_literal_chunk
 *      (core/src/gpu/dispatch.rs:57-301) + dispatch::scan_regex_chunk (dispatch.rs:331-380).
 * data may be host or device memory (data_mem); out likewise (out_mem).
 * On WS_ERR_MATCH_BUFFER_OVERFLOW *count holds the required capacity. */
ws_status ws_scan(ws_ctx* ctx, const ws_patterns* patterns, const uint8_t* data, uint64_t len,
                  uint32_t data_mem, uint32_t mode, const ws_tuning* tuning, ws_match* out,
                  uint64_t cap, uint32_t out_mem, uint64_t* count);

/* ---- async scan: GpuMatcher::scan's async contract (core/src/gpu/mod.rs:223-260: the
 * future resolves when the device's results are read back).
 * ws_scan_async starts ws_scan with the same arguments and returns once the call is
 * enqueued (overlapping mode on the device-sized pipeline; other modes and pipelines
 * finish before it returns). The input, output and pattern set must stay valid and
 * unchanged until ws_wait. One call per context may be outstanding: a second
 * ws_scan_async before ws_wait is WS_ERR_INVALID_ARGUMENT. Any other call on the
 * context first completes the outstanding one, whose result ws_wait still returns.
 * A non-OK status from ws_scan_async means nothing is outstanding. */
ws_status ws_scan_async(ws_ctx* ctx, const ws_patterns* patterns, const uint8_t* data, uint64_t len,
                        uint32_t data_mem, uint32_t mode, const ws_tuning* tuning, ws_match* out,
                        uint64_t cap, uint32_t out_mem);
/* 1 when ws_wait would return without waiting for the device (or nothing is outstanding),
 * 0 while the device is still working on the call. */
int ws_poll(ws_ctx* ctx);
/* Completes the outstanding ws_scan_async call: its status and *count exactly as ws_scan
 * would have returned them. WS_ERR_INVALID_ARGUMENT when nothing is outstanding. */
ws_status ws_wait(ws_ctx* ctx, uint64_t* count);

/* ---- allocation-free scan. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). ws_scan in overlapping mode for
 * device-resident data and a device-resident out array, with every workspace
 * in one fixed array per GPU that the library sets up once per device and
 * keeps until the process ends (20 MiB); the call itself allocates nothing.
 * Limits: len <= 16 MiB (WS_ERR_INPUT_TOO_LARGE beyond); at most 65,536
 * literal matches and 65,536 regex candidates per call; an overflow of these
 * fixed buffers (also: matches too clustered for its buckets) is
 * WS_ERR_MATCH_BUFFER_OVERFLOW with *count = 0 and no result (ws_scan then
 * gives one). A full `out` is WS_ERR_MATCH_BUFFER_OVERFLOW with the required
 * count. Every context on a GPU shares that GPU's array: calls that overlap
 * on one GPU race and must be serialised by the caller. */
ws_status ws_scan_noalloc(ws_ctx* ctx, const ws_patterns* patterns, const uint8_t* data, uint64_t len,
                          const ws_tuning* tuning, ws_match* out, uint64_t cap, uint64_t* count);

/* ---- batch: batch::scan_batch / scan_batch_gpu / scan_batch_cpu (core/src/batch.rs:170-363).
 * Every item is scanned on its own (no match crosses items). Item data is host memory
 * when data_mem is WS_MEM_HOST, else device pointers. One persistent filter launch;
 * overlapping-mode scans on the hash filter are device-sized (one synchronisation). */
ws_status ws_scan_batch(ws_ctx* ctx, const ws_patterns* patterns, const ws_item* items,
                        size_t n_items, uint32_t data_mem, uint32_t mode,
                        const ws_tuning* tuning, ws_tagged_match* out, uint64_t cap,
                        uint64_t* count);

/* ---- batch plans: the host planning of scan_batch done once -- batch::coalesce's
 * layout and planned_batches' grouping (core/src/batch.rs:65-104, 197-236) for
 * items in device memory: id order, each item's segment, work units largest
 * item first for every lane width, dense item positions and ids, all copied to
 * the device. The items must stay in place, unchanged in length, while the plan
 * is used. A plan may be scanned with any pattern set on its device, from any
 * context on that device. */
typedef struct ws_batch_plan ws_batch_plan;
ws_status ws_batch_plan_create(ws_ctx* ctx, const ws_item* items, size_t n_items, ws_batch_plan** out);
void ws_batch_plan_destroy(ws_batch_plan* plan);
/* scan_batch over a plan (core/src/batch.rs:170-195): results as ws_scan_batch,
 * into host or device memory (out_mem). Overlapping-mode scans on the hash
 * filter are device-sized (one synchronisation) and a repeated call (same
 * patterns, plan, output and tuning) replays a captured hipGraph; other modes
 * and variants take the checked pipeline. */
ws_status ws_scan_batch_planned(ws_ctx* ctx, const ws_patterns* patterns, const ws_batch_plan* plan,
                                uint32_t mode, const ws_tuning* tuning, ws_tagged_match* out, uint64_t cap,
                                uint32_t out_mem, uint64_t* count);

/* ---- stream: StreamPipeline::scan (core/src/pipeline.rs:86-153).
 * Scans in chunks of chunk_size bytes; each chunk reads `overlap` extra bytes and keeps
 * matches that start inside it. Regex non-overlap chains carry across chunks, so the
 * result equals one ws_scan over the whole input when overlap >= the longest match.
 * WS_MODE_STANDARD with literal reductions over more than one chunk needs overlap >= the
 * longest literal - 1 (WS_ERR_INVALID_ARGUMENT otherwise): the reductions are one greedy
 * pass over the whole input, so such a call returns ws_scan's whole-input result. */
ws_status ws_stream_scan(ws_ctx* ctx, const ws_patterns* patterns, const uint8_t* data,
                         uint64_t len, uint32_t data_mem, uint64_t chunk_size, uint64_t overlap,
                         uint32_t mode, const ws_tuning* tuning, ws_match* out, uint64_t cap,
                         uint32_t out_mem, uint64_t* count);

/* ---- HotSwapPatternSet (core/src/pattern/mod.rs:17-70): the active compiled set,
 * replaced atomically. A scan takes a reference to the set that is active when it
 * starts and keeps it to the end, so a swap never changes a scan in flight. */
typedef struct ws_hotswap ws_hotswap;

/* HotSwapPatternSet::new / from_arc: takes its own reference to `initial`. */
ws_status ws_hotswap_create(ws_patterns* initial, ws_hotswap** out);
void ws_hotswap_destroy(ws_hotswap* hs);
/* HotSwapPatternSet::swap_arc: 1 if `next` replaced the active set, 0 if it already
 * was the active set (Arc::ptr_eq); takes its own reference to `next`. */
int ws_hotswap_swap(ws_hotswap* hs, ws_patterns* next);
/* HotSwapPatternSet::scan: ws_scan with the active set. */
ws_status ws_hotswap_scan(ws_ctx* ctx, ws_hotswap* hs, const uint8_t* data, uint64_t len,
                          uint32_t data_mem, uint32_t mode, const ws_tuning* tuning, ws_match* out,
                          uint64_t cap, uint32_t out_mem, uint64_t* count);

/* ---- incremental stream: StreamScanner (core/src/stream.rs:14-135).
 * Each feed scans the carried overlap plus the new chunk (on the device) in `mode`,
 * reports the matches that end past the overlap, with offsets from the start of the
 * stream, and carries the window's last L bytes, L = the longest literal pattern
 * (regexes add nothing to L, as in core/src/pattern/compiler.rs:149,201). Offsets and
 * the carried window are limited to 4 GiB (WS_ERR_INPUT_TOO_LARGE). */
typedef struct ws_stream_scanner ws_stream_scanner;

/* StreamScanner::new (core/src/stream.rs:38-57); mode WS_MODE_STANDARD is the reference's. */
ws_status ws_stream_scanner_create(ws_ctx* ctx, const ws_patterns* patterns, uint32_t mode,
                                   ws_stream_scanner** out);
void ws_stream_scanner_destroy(ws_stream_scanner* scanner);
/* StreamScanner::feed (core/src/stream.rs:59-125). chunk is host or device memory
 * (data_mem); out is host memory. On WS_ERR_MATCH_BUFFER_OVERFLOW *count holds the
 * required capacity and the scanner has not advanced. */
ws_status ws_stream_scanner_feed(ws_stream_scanner* scanner, const uint8_t* chunk, uint64_t len,
                                 uint32_t data_mem, ws_match* out, uint64_t cap, uint64_t* count);
/* StreamScanner::finish (core/src/stream.rs:127-134): a feed of nothing. */
ws_status ws_stream_scanner_finish(ws_stream_scanner* scanner, ws_match* out, uint64_t cap,
                                   uint64_t* count);
/* Bytes fed so far (StreamScanner::processed_bytes). */
uint64_t ws_stream_scanner_processed(const ws_stream_scanner* scanner);

/* ---- compiled pattern index: CompiledPatternIndex (core/src/compiled_index/).
 * The WPSIDX01 little-endian format. Versions 1-3 are the reference's files:
 * they are read, checked as CompiledPatternIndex::load checks them, and rebuilt
 * from the literals and regex sources they store, as to_pattern_set does
 * (core/src/compiled_index/mod.rs:287-372); their regex-automata DFA sections
 * are checked for structure only. Version 3 is written for sets without
 * regexes (a file the reference loads). Version 5 is this library's default:
 * the version-3 sections, then each regex's anchored DFA and the factor
 * automaton, so that loading skips the subset construction and the automaton
 * build; version-4 files (an earlier layout of those sections) are read and
 * rebuilt from their sources. Versions 3 to 5 end in a CRC-32 (IEEE) of the
 * rest. The builders
 * write into `out` (host memory): with out == NULL they set *len to the size
 * and return WS_OK; with cap < the size they set *len and return
 * WS_ERR_INVALID_ARGUMENT. Malformed files are WS_ERR_PATTERN_COMPILATION_FAILED. */
#define WS_INDEX_VERSION 5u
enum { WS_INDEX_REFERENCE_V3 = 1u }; /* index_flags: write version 3 */

typedef struct ws_index_info {
  uint32_t version;
  uint32_t case_insensitive;
  uint32_t pattern_count;
  uint32_t literal_count;
  uint32_t regex_count;
  uint32_t hash_window_len;
  uint32_t prebuilt;          /* 1 for version 5: tables load without compilation */
  uint32_t reserved;
} ws_index_info;

/* CompiledPatternIndex::build (core/src/compiled_index/builder.rs:8-61) from
 * pattern sources, on the host (flags as for ws_patterns_compile). */
ws_status ws_index_build(const ws_pattern* patterns, size_t count, uint32_t flags,
                         uint32_t index_flags, uint8_t* out, uint64_t cap, uint64_t* len);
/* The same for a compiled set. */
ws_status ws_patterns_serialize(const ws_patterns* patterns, uint32_t index_flags, uint8_t* out,
                                uint64_t cap, uint64_t* len);
/* CompiledPatternIndex::load (core/src/compiled_index/mod.rs:58-61, 96-271) and
 * to_pattern_set, on the host only: checks `data`, reports its header and, when
 * `info` is not NULL, the facts of the set it holds (as ws_patterns_analyze). */
ws_status ws_index_inspect(const uint8_t* data, uint64_t len, ws_index_info* out,
                           ws_patterns_info* info);
/* load + to_pattern_set + device upload: a compiled set equal to compiling the
 * stored patterns (data is host memory). */
ws_status ws_index_load(ws_ctx* ctx, const uint8_t* data, uint64_t len, ws_patterns** out);
/* CompiledPatternIndex::load_from_file (core/src/compiled_index/mod.rs:76-94):
 * files above 512 MiB are WS_ERR_INPUT_TOO_LARGE. */
ws_status ws_index_load_file(ws_ctx* ctx, const char* path, ws_patterns** out);
/* CompiledPatternIndex::save_to_file (core/src/compiled_index/mod.rs:65-73). */
ws_status ws_index_save_file(const uint8_t* data, uint64_t len, const char* path);
/* CompiledPatternIndex::names (core/src/compiled_index/mod.rs:274-276): the name of
 * pattern `id` into buf (not NUL-terminated); *len = UINT64_MAX when it has none.
 * With cap < the length, *len holds it and the status is WS_ERR_INVALID_ARGUMENT. */
ws_status ws_patterns_name(const ws_patterns* patterns, uint32_t id, char* buf, uint64_t cap,
                           uint64_t* len);

/* ---- kernel routing: AutoMatcher (core/src/router.rs:205-291, core/src/config.rs:4-35).
 * The reference routes on one input-size threshold. Here an exhaustive autotuner
 * times every kernel configuration, and a small learned cost model (FP32 MLP,
 * trained on the GPU from the autotuner's timings) predicts the fastest one.
 * Routing never changes results: every configuration returns the same matches. */
#define WS_ROUTER_FEATURES 7
#define WS_ROUTER_CONFIGS 16

typedef struct ws_router ws_router;

/* Configuration `index` (< WS_ROUTER_CONFIGS) of the routing table. */
ws_status ws_config_tuning(uint32_t index, ws_tuning* out);
/* Router features of a request: log2 bytes, log2 patterns, log2 factor-DFA states,
 * byte classes, literal fraction, min literal length, log2(1 + matches per MiB);
 * out[7] is padding (0). */
ws_status ws_router_features(const ws_patterns_info* info, uint64_t bytes, float matches_per_mib,
                             float out[8]);
/* Exhaustive autotune: times each supported configuration of the routing table
 * whose bit is set in `configs` (0 = every one) `repeats` times on the request
 * (device or host data) and writes the best time per configuration (ms; 0 where
 * not timed), mask[i] = 1 for timed ones, and the fastest index. */
ws_status ws_autotune(ws_ctx* ctx, const ws_patterns* patterns, const uint8_t* data, uint64_t len,
                      uint32_t data_mem, uint32_t mode, uint32_t repeats, uint32_t configs,
                      float* times_ms, float* mask, uint32_t* best);

ws_status ws_router_create(int device_ordinal, uint32_t seed, ws_router** out);
void ws_router_destroy(ws_router* router);
const char* ws_router_last_error(void);
uint32_t ws_router_param_count(void);
ws_status ws_router_get_params(ws_router* router, float* out, uint32_t n);
ws_status ws_router_set_params(ws_router* router, const float* in, uint32_t n);
ws_status ws_router_set_normalization(ws_router* router, const float* xmean, const float* xstd,
                                      float ymean, float ystd);
ws_status ws_router_get_normalization(ws_router* router, float* xmean, float* xstd, float* ymean,
                                      float* ystd);
/* Trains on raw features x[n][8] and log-ms targets y[n][WS_ROUTER_CONFIGS] (mask, same shape: 1
 * where measured; a request may be timed on part of the table). */
ws_status ws_router_train(ws_router* router, const float* x, const float* y, const float* mask,
                          uint32_t n, uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                          float* final_loss);
/* Predicted ms per configuration (pred_ms[n][WS_ROUTER_CONFIGS], optional) and the fastest
 * configuration allowed by mask[n][WS_ROUTER_CONFIGS]. */
ws_status ws_router_predict(ws_router* router, const float* x, uint32_t n, const float* mask,
                            float* pred_ms, uint32_t* choice);
/* Raw outputs / loss gradient on standardized inputs (for checking against a CPU reference). */
ws_status ws_router_forward(ws_router* router, const float* x, uint32_t n, float* out);
ws_status ws_router_gradient(ws_router* router, const float* x, const float* y, const float* mask,
                             uint32_t n, float scale, float* grad_out);

#ifdef __cplusplus
}
#endif
#endif /* WARPSTATE_AMD_H */
