/* f5c's two ABEA entry points called from plain C through libabea, as the
 * C shim a maintainer adds to f5c would call them (align() per read,
 * GPU/f5cmisc.h:48-50; align_cuda() over a batch, GPU/f5cmisc.h:99-102).
 *
 *   c_dropin <in.bin> <out.bin>
 *
 * in.bin:  int32 n_reads; the model (4096 x 12 B); then per read int32 len,
 *          int32 n_events, scalings (16 B), the bases (len bytes), the event
 *          records (n_events x 24 B).
 * out.bin: for every read, the pairs abea_align() returned (int32 count, then
 *          the pairs); then for every read the pairs abea_align_db() returned
 *          (the same form).
 *
 * The layout checks restate the x86-64 sizes and offsets of f5c's types that
 * a caller casts between (event_t, event_table, model_t with CACHED_LOG,
 * scalings_t, AlignedPair). */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "abea.h"

_Static_assert(sizeof(abea_event_t) == 24, "event_t is 24 bytes");
_Static_assert(offsetof(abea_event_t, start) == 0, "event_t.start");
_Static_assert(offsetof(abea_event_t, length) == 8, "event_t.length");
_Static_assert(offsetof(abea_event_t, mean) == 12, "event_t.mean");
_Static_assert(offsetof(abea_event_t, stdv) == 16, "event_t.stdv");
_Static_assert(sizeof(abea_event_table) == 32, "event_table is 32 bytes");
_Static_assert(offsetof(abea_event_table, event) == 24, "event_table.event");
_Static_assert(sizeof(abea_model_t) == 12, "model_t is 12 bytes");
_Static_assert(sizeof(abea_scalings_t) == 16, "scalings_t is 16 bytes");
_Static_assert(sizeof(abea_aligned_pair_t) == 8, "AlignedPair is 8 bytes");

static void die(const char *what) {
  fprintf(stderr, "c_dropin: %s\n", what);
  exit(2);
}

static void read_exact(FILE *f, void *p, size_t n) {
  if (n && fread(p, 1, n, f) != n) die("short input file");
}

static void write_exact(FILE *f, const void *p, size_t n) {
  if (n && fwrite(p, 1, n, f) != n) die("cannot write the output file");
}

int main(int argc, char **argv) {
  if (argc != 3) die("usage: c_dropin <in.bin> <out.bin>");
  FILE *in = fopen(argv[1], "rb");
  if (!in) die("cannot open the input file");
  int32_t n = 0;
  read_exact(in, &n, sizeof n);
  if (n <= 0 || n > 1000000) die("bad read count");
  abea_model_t *model = malloc(sizeof(abea_model_t) * ABEA_NUM_KMER);
  char **read = calloc((size_t)n, sizeof(char *));
  int32_t *read_len = calloc((size_t)n, sizeof(int32_t));
  abea_event_table *et = calloc((size_t)n, sizeof(abea_event_table));
  abea_scalings_t *scal = calloc((size_t)n, sizeof(abea_scalings_t));
  if (!model || !read || !read_len || !et || !scal) die("out of memory");
  read_exact(in, model, sizeof(abea_model_t) * ABEA_NUM_KMER);
  for (int32_t i = 0; i < n; ++i) {
    int32_t n_e = 0;
    read_exact(in, &read_len[i], sizeof(int32_t));
    read_exact(in, &n_e, sizeof(int32_t));
    if (read_len[i] < 0 || n_e < 0) die("bad read header");
    read_exact(in, &scal[i], sizeof(abea_scalings_t));
    read[i] = malloc((size_t)read_len[i] + 1);
    et[i].event = malloc(sizeof(abea_event_t) * ((size_t)n_e + 1));
    if (!read[i] || !et[i].event) die("out of memory");
    read_exact(in, read[i], (size_t)read_len[i]);
    read[i][read_len[i]] = '\0';  /* f5c's sequences are NUL-terminated */
    read_exact(in, et[i].event, sizeof(abea_event_t) * (size_t)n_e);
    et[i].n = (size_t)n_e;
    et[i].start = 0;
    et[i].end = (size_t)n_e;
  }
  fclose(in);

  FILE *out = fopen(argv[2], "wb");
  if (!out) die("cannot open the output file");
  /* align(): one read per call, into f5c's 2 * n_events buffer */
  for (int32_t i = 0; i < n; ++i) {
    abea_aligned_pair_t *out_2 = malloc(sizeof(abea_aligned_pair_t) * (2 * et[i].n + 1));
    if (!out_2) die("out of memory");
    const int32_t np = abea_align(out_2, read[i], read_len[i], et[i], model, scal[i], 4000.0f);
    if (np < 0) {
      fprintf(stderr, "c_dropin: abea_align: %s\n", abea_last_error());
      return 1;
    }
    write_exact(out, &np, sizeof np);
    write_exact(out, out_2, sizeof(abea_aligned_pair_t) * (size_t)np);
    free(out_2);
  }
  /* align_cuda(): the batch, callee-malloc'd pairs per read */
  abea_aligned_pair_t **pairs = calloc((size_t)n, sizeof(abea_aligned_pair_t *));
  int32_t *n_pairs = calloc((size_t)n, sizeof(int32_t));
  if (!pairs || !n_pairs) die("out of memory");
  const abea_status_t st =
      abea_align_db(n, read, read_len, et, model, scal, pairs, n_pairs);
  if (st != ABEA_OK) {
    fprintf(stderr, "c_dropin: abea_align_db: %s\n", abea_last_error());
    return 1;
  }
  for (int32_t i = 0; i < n; ++i) {
    write_exact(out, &n_pairs[i], sizeof(int32_t));
    write_exact(out, pairs[i], sizeof(abea_aligned_pair_t) * (size_t)n_pairs[i]);
    free(pairs[i]);
  }
  fclose(out);
  for (int32_t i = 0; i < n; ++i) {
    free(read[i]);
    free(et[i].event);
  }
  free(pairs);
  free(n_pairs);
  free(read);
  free(read_len);
  free(et);
  free(scal);
  free(model);
  return 0;
}
