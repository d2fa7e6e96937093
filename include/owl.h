/* owl.h -- C-ABI of the modular-squaring engine (one PRP/LL iteration mod 2^E - 1
 * as an FP64 IBDWT on gfx950).
 *
 * Each entry point wraps a method of the reference's host class Gpu
 * (src/Gpu.h:213-300) or FFTConfig (src/FFTConfig.h:69-87), cited per function.
 * Plain pointers and sizes only. Returns 0 on success, non-zero on failure with
 * the message in owl_last_error() (declared in owl_hip.h). Buffers are named by
 * index: 0 = data, 1 = aux, 2 = check, 3 = base (src/Gpu.h:134-140).
 */
#ifndef OWL_H
#define OWL_H

#include <stddef.h>
#include <stdint.h>

#include "owl_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct owl_gpu_s* owl_gpu;

typedef struct {
  uint32_t E, N, W, H, R, nb, rowThreads, colThreads;
  uint32_t maxExp;
  int longCarry;
  char spec[64];     /* reference FFT spec, e.g. "1K:15:256:202" */
  char rowKernel[32], colKernel[32];   /* the row / column pass kernels in use */
} owl_gpu_info_t;

typedef struct {
  uint64_t res64;
  uint32_t k, nErrors, nChecks;
  int isPrime;
  double maxRoe, secsPerIt;
} owl_prp_result_t;

typedef struct {
  uint64_t res64;
  uint32_t k;
  int isPrime;
  double secsPerIt;
} owl_ll_result_t;

/* FFTConfig::bestFit (src/FFTConfig.cpp:235-263). spec may be "" or NULL: then the first FFT of the
 * loaded tune list whose maxExp covers E, else the first shape (variants 101, 202) that does.
 * maxExp is this engine's: the reference's bits-per-word limit (src/fftbpw.h), lowered for the shapes
 * where this engine's round-off at that limit is above the reference's criterion (Gumbel Z = 28). */
int owl_fft_best_fit(uint32_t E, const char* spec, char* outSpec, size_t outLen, uint32_t* N, uint32_t* maxExp);
/* The same selection with the reference's limits and no tune list: the FFT the reference picks. */
int owl_fft_best_fit_ref(uint32_t E, const char* spec, char* outSpec, size_t outLen, uint32_t* N, uint32_t* maxExp);
/* The reference's tune.txt (TuneEntry, src/TuneEntry.cpp:7-84): a process-wide list of FFTs ordered by
 * cost (microseconds per iteration) with maxExp increasing, read by bestFit when no spec is given
 * (owl_fft_best_fit, owl_gpu_create, batches, the server). Lines "cost spec [# maxExp]"; saved as
 * "%6.1f %14s # %u". load replaces the list (an unordered file is an error); add inserts an entry
 * unless an entry at most as costly covers at least its maxExp, removing the costlier entries it
 * covers (*updated = whether the list changed); entry i of the list: *found = 0 past the end. */
int owl_fft_tune_load(const char* path);
int owl_fft_tune_save(const char* path);
int owl_fft_tune_clear(void);
int owl_fft_tune_add(double cost, const char* spec, int* updated);
int owl_fft_tune_entry(uint32_t i, double* cost, char* spec, size_t specLen, uint32_t* maxExp, int* found);

/* Gpu::Gpu / Gpu::make (src/Gpu.cpp:380-564). lineBlock 0 = automatic.
 * flags: OWL_LONG_CARRY = full carry every iteration (the reference's -carry long);
 *        OWL_GENERIC = runtime-plan kernels even where a shape-specialized pass exists;
 *        OWL_COL_ONE_REGION = the column pass variant that takes the two columns through one
 *        LDS region in turn (pairing in registers) where the two-region pass exists;
 *        OWL_COL_ONE_PAIR = the two-region column pass with one column pair per group where
 *        the variant with two pairs per group would be chosen. */
enum { OWL_LONG_CARRY = 1, OWL_GENERIC = 2, OWL_COL_ONE_REGION = 4, OWL_COL_ONE_PAIR = 8 };
int owl_gpu_create(uint32_t E, const char* spec, int flags, uint32_t lineBlock, owl_gpu* gpu);
int owl_gpu_destroy(owl_gpu gpu);
int owl_gpu_info(owl_gpu gpu, owl_gpu_info_t* info);

/* Gpu::writeIn / readData (src/Gpu.cpp:801-811, 707-709): packed residue words, (E-1)/32+1 u32.
 * read: *isZero = 1 and no words written when the residue is all zero. */
int owl_gpu_write_words(owl_gpu gpu, int buf, const uint32_t* words, uint32_t nWords);
int owl_gpu_read_words(owl_gpu gpu, int buf, uint32_t* words, uint32_t nWords, int* isZero);
/* Raw balanced digits in sequential word order, N int32 (src/Gpu.cpp:677-705, 801-811). */
int owl_gpu_write_digits(owl_gpu gpu, int buf, const int32_t* digits, uint32_t n);
int owl_gpu_read_digits(owl_gpu gpu, int buf, int32_t* digits, uint32_t n);
/* No counterpart in the reference: owl_gpu_read_digits without any device allocation, for callers
 * that cannot allocate (code under graph capture, serving loops, harnesses that forbid it). The
 * digits are staged in a static device array of OWL_NOALLOC_MAX_WORDS = 2^23 int32 (32 MiB, held
 * per device until the process ends); N above that is rejected with an error. Calls on one device
 * share that array: overlapping calls on one device race. */
int owl_gpu_read_digits_noalloc(owl_gpu gpu, int buf, int32_t* digits, uint32_t n);

/* Gpu::dataResidue / bufResidue (src/Gpu.cpp:955-975). */
int owl_gpu_res64(owl_gpu gpu, int buf, uint64_t* res);

/* Gpu::squareLoop (src/Gpu.cpp:936-945): iterations [from, to); mul3Last multiplies the last by 3;
 * ll selects x -> x^2 - 2. */
int owl_gpu_square_loop(owl_gpu gpu, int out, int in, uint32_t from, uint32_t to, int mul3Last, int ll);
/* Gpu::modMul (src/Gpu.cpp:729-733): ioA <- ioA * inB (* 3 if mul3). */
int owl_gpu_mod_mul(owl_gpu gpu, int ioA, int inB, int mul3);
/* Gpu::doCheck (src/Gpu.cpp:767-771) and Gpu::isEqual (src/Gpu.cpp:947-953). */
int owl_gpu_do_check(owl_gpu gpu, uint32_t blockSize, int* ok);
int owl_gpu_is_equal(owl_gpu gpu, int a, int b, int* equal);
/* Buffer copy (bufAux << bufData, src/Gpu.cpp:741-742). */
int owl_gpu_copy(owl_gpu gpu, int dst, int src);
/* Gpu::readROE (src/Gpu.cpp:624-637): statistics of the max round-off per iteration since the last read. */
int owl_gpu_read_roe(owl_gpu gpu, uint32_t* n, double* maxRoe, double* mean, double* sd);

/* Gpu::measureROE (src/Gpu.cpp:1268-1335) with RoeInfo's Gumbel fit (src/Gpu.h:62-81): round-off of
 * iters (<= 100,000) squarings of a PRP-3 run from 3 after warmup unmeasured ones, one sample per
 * squaring; *z = (0.5 - mu) / beta with beta = sd*sqrt(6)/pi, mu = mean - 0.5772*beta (the reference
 * fits its bits-per-word limits at Z = 28, src/tune.cpp:106-172). */
int owl_gpu_measure_roe(owl_gpu gpu, uint32_t iters, uint32_t warmup, uint32_t* n, double* maxRoe, double* mean,
                        double* sd, double* z);
/* Gpu::isPrimePRP (src/Gpu.cpp:1400-1584) without savefiles or proofs (owl_gpu_prp_saved and
 * owl_gpu_prp_proof below add them): iters PRP-3 iterations from x0 = 3 with a Gerbicz check every
 * checkStep (multiple of blockSize). A full test (iters == E) runs on to the next multiple of
 * blockSize and is checked there; res64 and isPrime are those of x_E. res64 after every resStep
 * iterations goes to res[0..resCap). */
int owl_gpu_prp(owl_gpu gpu, uint32_t iters, uint32_t blockSize, uint32_t checkStep, uint64_t* res, uint32_t resStep,
                uint32_t resCap, owl_prp_result_t* out);
/* Failure-handling tests (the reference tests its recovery by hand, src/Gpu.cpp:1561-1575):
   the next PRP run flips one data bit right after iteration k, once. 0 disarms. */
int owl_gpu_arm_fault(owl_gpu gpu, uint32_t k);
/* Gpu::isPrimeLL (src/Gpu.cpp:1586-1663) without savefiles (owl_gpu_ll_saved adds them): iters LL
 * iterations from x0 = 4. */
int owl_gpu_ll(owl_gpu gpu, uint32_t iters, uint64_t* res, uint32_t resStep, uint32_t resCap, owl_ll_result_t* out);

/* Savefiles in the reference's formats (Saver<PRPState> / Saver<LLState>, src/Saver.cpp:17-30,
 * 89-143, 160-298; File::writeChecked / readChecked, src/File.h:213-231; CycleFile,
 * src/CycleFile.cpp:6-21). owl_gpu_prp_saved / owl_gpu_ll_saved are isPrimePRP / isPrimeLL
 * with a Saver on directory dir: the run starts from the most recent savefile in
 * dir/<E>/ (PRP) or dir/ll-<E>/ (LL), or from x0 when there is none (a PRP file is accepted
 * when the res64 of x_k = 3 * check^(2^B - 1) matches its header, else the one before it is
 * tried, src/Gpu.cpp:1169-1191), writes <E>-<k>.prp at every verified Gerbicz check point
 * below E (src/Gpu.cpp:1545-1551) or <E>-<k>.ll every saveStep iterations and at the end
 * (the reference saves LL every 10,000, src/Gpu.cpp:1619-1646), and keeps the newest
 * nSavefiles. res entries before the starting k are 0. */
int owl_gpu_prp_saved(owl_gpu gpu, const char* dir, uint32_t nSavefiles, uint32_t iters, uint32_t blockSize,
                      uint32_t checkStep, uint64_t* res, uint32_t resStep, uint32_t resCap, owl_prp_result_t* out);
int owl_gpu_ll_saved(owl_gpu gpu, const char* dir, uint32_t nSavefiles, uint32_t saveStep, uint32_t iters, uint64_t* res,
                     uint32_t resStep, uint32_t resCap, owl_ll_result_t* out);
/* One savefile: the PRP v13 / LL v13 header line, the CRC-32 of the words, then the
 * (E-1)/32+1 words (check for PRP, data for LL), written through "<path>.new" and a rename.
 * Reading also takes PRP v12 and LL v1 files; a bad header, short file or CRC mismatch is an
 * error. The read functions fill the words when words != NULL (cap >= (E-1)/32+1). */
int owl_savefile_write_prp(const char* path, uint32_t E, uint32_t k, uint32_t blockSize, uint64_t res64,
                           uint32_t nErrors, double elapsed, const uint32_t* check, uint32_t nWords);
int owl_savefile_read_prp(const char* path, uint32_t* E, uint32_t* k, uint32_t* blockSize, uint64_t* res64,
                          uint32_t* nErrors, double* elapsed, uint32_t* check, uint32_t cap);
int owl_savefile_write_ll(const char* path, uint32_t E, uint32_t k, double elapsed, const uint32_t* data,
                          uint32_t nWords);
int owl_savefile_read_ll(const char* path, uint32_t* E, uint32_t* k, double* elapsed, uint32_t* data, uint32_t cap);
/* Worktodo lines (src/Worktodo.cpp:25-86): kind OWL_TASK_PRP ("PRP=", "PRPDC="), OWL_TASK_LL ("Test=",
 * "DoubleCheck="), OWL_TASK_CERT ("Cert=AID,1,2,E,-1,squarings") or OWL_TASK_NONE (comments, unknown
 * kinds, exponents <= 1000); aid gets the 32-hex-digit assignment id or "". */
enum { OWL_TASK_NONE = 0, OWL_TASK_PRP = 1, OWL_TASK_LL = 2, OWL_TASK_CERT = 3 };
int owl_worktodo_parse(const char* line, int* kind, uint32_t* exponent, char* aid, size_t aidLen, uint32_t* squarings);
/* bestTask (src/Worktodo.cpp:96-106): the index of the smallest CERT, else of the smallest exponent, among
 * n lines; -1 when none is a task. */
int owl_worktodo_best(const char* const* lines, uint32_t n, int32_t* index);
/* The PRP-3 type-1 residue reported at k = E (src/Gpu.cpp:1120-1148, 1495-1504): q = x * 9^-1 mod
 * 2^E - 1 of the (E-1)/32+1 words x; its res64 and res2048 (the 64 low words as 512 hex digits, most
 * significant first; len >= 513). */
int owl_prp_type1(const uint32_t* words, uint32_t nWords, uint32_t E, uint64_t* res64, char* res2048, size_t len);
/* One results JSON line (src/Task.cpp:86-170) for kind OWL_TASK_PRP ("PRP-3": res64, res2048,
 * residue-type 1, errors.gerbicz, fft-length) or OWL_TASK_LL (res64, fft-length, shift-count,
 * error-code), with the program block (programName / programVersion; NULL = this engine's), user,
 * aid, uid (NULL or "" omitted) and a UTC timestamp. */
int owl_result_json(int kind, uint32_t E, int isPrime, uint64_t res64, const char* res2048, uint32_t fftLength,
                    uint32_t nErrors, const char* aid, const char* user, const char* uid, const char* programName,
                    const char* programVersion, char* out, size_t outLen);

/* Results lines of every kind (Task::writeResultPRP / writeResultLL / writeResultCERT,
 * src/Task.cpp:148-197): kind OWL_TASK_PRP, OWL_TASK_LL or OWL_TASK_CERT. PRP adds the
 * "proof" object {version 1, power, hashsize 64, md5} after fft-length when proofPower > 0
 * (proofMd5: owl_proof_md5 of the proof file). A CERT line has no status and carries worktype
 * "Cert", exponent, sha3-hash (64 hex digits, as owl_gpu_cert writes it), squarings,
 * fft-length, shift-count 0 and error-code. NULL strings count as "". */
typedef struct {
  int kind;
  uint32_t exponent;
  int isPrime;
  uint64_t res64;
  const char* res2048;
  uint32_t fftLength, nErrors;
  uint32_t proofPower;
  const char* proofMd5;
  const char* sha3Hash;
  uint32_t squarings;
  const char *aid, *user, *uid, *programName, *programVersion;
} owl_result_t;
int owl_result_json_ex(const owl_result_t* result, char* out, size_t outLen);
/* MD5 of a file as 32 lowercase hex digits (proof::fileHash, src/Proof.cpp:32-39); len >= 33. */
int owl_proof_md5(const char* path, char* md5Hex32, size_t len);

/* Gpu::isCERT (src/Gpu.cpp:1665-1725): reads the start value from certPath ((E-1)/8+1
 * little-endian bytes, the reference's M<E>.cert), squares it `squarings` times, and writes
 * the SHA3-256 of the (E-1)/8+1 bytes of the result as the reference's results line reports
 * it (its four little-endian 64-bit words from the last to the first, 64 hex digits; len >= 65)
 * and the result's res64. removeFile deletes certPath when done, as the reference does. */
int owl_gpu_cert(owl_gpu gpu, const char* certPath, uint32_t squarings, int removeFile, char* sha3Hex, size_t len,
                 uint64_t* res64);

/* PRP proofs (Proof / ProofSet, src/Proof.cpp:20-297; Gpu::expMul / expExp2, src/Gpu.cpp:813-889).
 * owl_gpu_prp_proof runs the full PRP-3 test (E iterations, Gerbicz checks as owl_gpu_prp) keeping the
 * residues of the 2^power proof points (owl_proof_points) under residueDir, builds the proof and
 * writes it to proofPath ("PRP PROOF
VERSION=2
HASHSIZE=64
POWER=p
NUMBER=M<E>
", then B and the
 * middles as (E-1)/8+1 little-endian bytes each), then verifies it as the reference does before
 * keeping one (an invalid proof is an error). owl_gpu_proof_verify checks a proof file:
 * *ok = 1 when valid, *isPrime = 1 when its B is 9. owl_gpu_exp_mul: out = A^h * B (B squared first
 * when squareB), all nWords = (E-1)/32+1 words. power: 1..12. */
int owl_proof_points(uint32_t E, uint32_t power, uint32_t* points, uint32_t cap, uint32_t* n);
int owl_gpu_prp_proof(owl_gpu gpu, const char* residueDir, uint32_t power, uint32_t blockSize, uint32_t checkStep,
                      const char* proofPath, owl_prp_result_t* out);
int owl_gpu_proof_verify(owl_gpu gpu, const char* proofPath, int* ok, int* isPrime);
int owl_gpu_exp_mul(owl_gpu gpu, const uint32_t* A, uint64_t h, const uint32_t* B, int squareB, uint32_t* out,
                    uint32_t nWords);
/* SHA3-256 (FIPS 202) of size bytes, the proof's hash (src/sha3.cpp): 32 bytes to out. */
int owl_sha3_256(const void* data, size_t size, uint8_t* out);

/* root1 (src/TrigBufCache.cpp:22-52): the table entry exp(-2 pi i k / n) = (*c, *s) that every
 * twiddle table of the engine holds, computed in long double from an exactly reduced angle and
 * rounded once to double. */
int owl_trig_root1(uint32_t n, uint64_t k, double* c, double* s);

/* crc32 (src/common.cpp:26-39): CRC-32/IEEE of size bytes. */
int owl_crc32(const void* data, size_t size, uint32_t* crc);

/* Gpu::timePRP (src/Gpu.cpp:1337-1398): microseconds per PRP iteration incl. Gerbicz updates. */
int owl_gpu_time_prp(owl_gpu gpu, uint32_t iters, uint32_t warmup, uint32_t blockSize, double* usPerIt);
/* Steady-state squarings only (no reference counterpart): us per iteration, optionally replayed from a
 * captured graph. */
int owl_gpu_time_squares(owl_gpu gpu, uint32_t iters, uint32_t warmup, int useGraph, double* usPerIt);
/* Per-kernel average durations over iters steady-state iterations, timed with events on the engine's
 * queue around each launch (no reference counterpart; the reference's -time profile is src/Gpu.cpp:773-799). */
int owl_gpu_time_kernels(owl_gpu gpu, uint32_t iters, double* usCol, double* usRow);

/* Device-to-device copy of the FP64 state (8*N bytes read, 8*N written: the same bytes as one
 * pass), timed with events over iters copies: the streaming ceiling of the box and the
 * calibration kernel for HBM traffic counters (no reference counterpart). */
int owl_gpu_time_copy(owl_gpu gpu, uint32_t iters, double* usPerCopy);

/* Batches of exponents (no reference counterpart: the reference runs one exponent per Gpu,
 * with at most four worker threads on one device, src/main.cpp:110-117). Jobs of equal FFT
 * geometry share every launch; groups of different geometry run on their own queues.
 * spec "" gives every exponent the reference's FFT choice. res is n rows of resCap res64. */
typedef struct owl_batch_s* owl_batch;
int owl_batch_create(const uint32_t* exponents, uint32_t n, const char* spec, int flags, owl_batch* batch);
int owl_batch_destroy(owl_batch batch);
int owl_batch_info(owl_batch batch, uint32_t* nJobs, uint32_t* nGroups, uint64_t* totalWords);
int owl_batch_job_info(owl_batch batch, uint32_t job, owl_gpu_info_t* info);
int owl_batch_prp(owl_batch batch, uint32_t iters, uint32_t blockSize, uint32_t checkStep, uint64_t* res,
                  uint32_t resStep, uint32_t resCap, owl_prp_result_t* out);
int owl_batch_arm_fault(owl_batch batch, uint32_t job, uint32_t k);
int owl_batch_ll(owl_batch batch, uint32_t iters, uint64_t* res, uint32_t resStep, uint32_t resCap,
                 owl_ll_result_t* out);
/* microseconds per batch iteration (every job squared once), optionally from captured graphs */
int owl_batch_time_squares(owl_batch batch, uint32_t iters, uint32_t warmup, int useGraph, double* usPerIter);

/* Serving (no reference counterpart: the reference has no request loop; each job follows
 * Gpu::isPrimePRP / isPrimeLL, src/Gpu.cpp:1400-1663). Requests are admitted while their
 * device footprint fits memBudget (0 = 3/4 of free memory) and fewer than maxJobs run; every
 * owl_server_step admits, runs one round of blockSize iterations on every active job (jobs of
 * one FFT geometry and kind share launches), verifies PRP jobs due for a Gerbicz check
 * (every checkStep, and past the end) and retires finished ones. */
typedef struct owl_server_s* owl_server;
typedef struct {
  uint32_t id, E, iters;
  int ll;
  uint64_t res64;   /* res64 of x_iters */
  int isPrime;      /* iters == E (PRP, x == 9) or E - 2 (LL, x == 0) */
  uint32_t nErrors, nChecks;
  double maxRoe;
  double tSubmit, tAdmit, tDone; /* seconds on the server's clock (owl_server_now) */
  int status;       /* 0 done, 1 failed (repeated Gerbicz errors), 2 rejected */
  /* a PRP request run to iters == E: the type-1 residue x_E * 9^-1 that results lines report
   * (res64 and the 64 low words as 512 hex digits, as owl_prp_type1); 0 and "" otherwise */
  uint64_t res64Type1;
  char res2048[516];
} owl_serve_result_t;
typedef struct {
  uint32_t waiting, active, groups, done, rounds;
  uint64_t bytesInUse, budget;
  uint64_t iterations;      /* requested squarings completed, all jobs */
  uint64_t wordIterations;  /* the same weighted by FFT size N */
} owl_serve_info_t;
int owl_server_create(uint64_t memBudget, uint32_t maxJobs, uint32_t blockSize, uint32_t checkStep, const char* spec,
                      int flags, owl_server* server);
int owl_server_destroy(owl_server server);
int owl_server_submit(owl_server server, uint32_t E, int ll, uint32_t iters, uint32_t* id);
int owl_server_step(owl_server server, uint32_t* nDone);
/* Failure-handling tests (as owl_gpu_arm_fault): request id, a PRP request still waiting, gets one
 * data bit flipped at the end of the round that passes iteration k; its next Gerbicz check fails
 * and the job reloads its last verified state (src/Gpu.cpp:1561-1575). */
int owl_server_arm_fault(owl_server server, uint32_t id, uint32_t k);
/* up to cap finished results (oldest first); *n is how many were written */
int owl_server_take(owl_server server, owl_serve_result_t* out, uint32_t cap, uint32_t* n);
int owl_server_info(owl_server server, owl_serve_info_t* info);
int owl_server_now(owl_server server, double* secs);

/* Autotuning (in place of the reference's tune.cpp / tune.txt, src/tune.cpp:328-738,
 * src/TuneEntry.cpp:7-84, which choose the FFT length; here the length stays bestFit's and
 * the tuner chooses how the engine runs it): every W x H split, line block and pass
 * variant for exponent E at bestFit(E, spec), timed over iters graph-replayed squarings,
 * fastest first. owl_tune_set / load record choices per FFT length N; owl_gpu_create,
 * owl_batch_create and the server use them when their flags and lineBlock choose nothing.
 * Tune file lines: "N width lineBlock flags us". */
typedef struct {
  uint32_t width, lineBlock;
  int flags;          /* OWL_GENERIC | OWL_COL_ONE_REGION | OWL_COL_ONE_PAIR */
  double usPerIter;
  uint32_t W, H, R;
  int longCarry;
  char rowKernel[32], colKernel[32];
} owl_tune_entry_t;
int owl_tune_run(uint32_t E, const char* spec, uint32_t iters, owl_tune_entry_t* out, uint32_t cap, uint32_t* n);
/* Batch tuning (no reference counterpart): for each FFT length among the n exponents, time a batch of
 * them under every W x H split over iters graph-replayed batch iterations (usPerIter = microseconds per
 * batch iteration) and record the fastest split in a table that owl_batch_create consults when its flags
 * choose nothing (single jobs keep the owl_tune_* table). Results grouped by length (N = 2 W H),
 * fastest first. */
int owl_tune_batch(const uint32_t* exponents, uint32_t n, const char* spec, uint32_t iters, owl_tune_entry_t* out,
                   uint32_t cap, uint32_t* nOut);
int owl_tune_batch_clear(void);
int owl_tune_set(uint32_t N, uint32_t width, uint32_t lineBlock, int flags, double usPerIter);
/* *found = 0 when no choice is recorded for N */
int owl_tune_get(uint32_t N, uint32_t* width, uint32_t* lineBlock, int* flags, double* usPerIter, int* found);
int owl_tune_clear(void);
int owl_tune_load(const char* path);
int owl_tune_save(const char* path);

/* Kernel-level access for numerics tests (no reference counterpart; the reference's
 * closest tools are its -verbose self-tests, src/Gpu.cpp:1059-1093).
 * owl_gpu_freq_io: read (write = 0) or write the FP64 state buffer idx (0 = state,
 * 2 = second-operand spectrum) as 2*N/2 doubles (re, im) in the engine's HBM layout.
 * owl_gpu_pass: run one pass on the queue and wait:
 *   OWL_PASS_LEADIN   words[arg] -> state (weight + forward row FFT)
 *   OWL_PASS_SQUARE   state <- column pass with the squaring
 *   OWL_PASS_FUSED    state <- row pass with the fused carry (arg = 1: LL)
 *   OWL_PASS_CARRYOUT state -> words[arg], fully carried
 *   (both measure their round-off into the next slot that owl_gpu_read_roe reports) 
 *   OWL_PASS_SPECTRUM spectrum buffer <- forward column transform of itself
 *   OWL_PASS_MUL      state <- column pass multiplying by the spectrum buffer */
enum { OWL_PASS_LEADIN = 0, OWL_PASS_SQUARE = 1, OWL_PASS_FUSED = 2, OWL_PASS_CARRYOUT = 3, OWL_PASS_SPECTRUM = 4,
       OWL_PASS_MUL = 5 };
int owl_gpu_freq_io(owl_gpu gpu, int idx, double* host, uint32_t nDoubles, int write);
int owl_gpu_pass(owl_gpu gpu, int pass, int arg);

#ifdef __cplusplus
}
#endif

#endif
