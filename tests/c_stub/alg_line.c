/* The C integration stub: align_benchmark's per-pair call sequence on the
 * library's WFA2-style surface.  Aligns argv[1] (pattern) against argv[2]
 * (text) with WFA2's default penalties and prints the pair's .alg line
 * ("score<TAB>CIGAR", benchmark_print_output_lite).
 *
 *   gcc -I<repo>/include alg_line.c -L<repo>/biwfa_gpu_pii2025_amd -lbiwfa_mi355x \
 *       -Wl,-rpath,<repo>/biwfa_gpu_pii2025_amd -o alg_line
 */
#include <stdio.h>
#include <string.h>

#include "biwfa_mi355x.h"

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s PATTERN TEXT\n", argv[0]);
    return 2;
  }
  const char* pattern = argv[1];
  const char* text = argv[2];
  const int plen = (int)strlen(pattern), tlen = (int)strlen(text);
  /* was: wavefront_aligner_attr_t attributes = wavefront_aligner_attr_default; ...
   *      wavefront_aligner_t* const wf_aligner = wavefront_aligner_new(&attributes); */
  bw_attr_t attr;
  bw_attr_default(&attr);
  attr.penalties = (bw_affine_penalties_t){.match = 0, .mismatch = 4, .gap_opening = 6, .gap_extension = 2};
  bw_aligner_t* const al = bw_aligner_new(&attr);
  if (al == NULL) {
    fprintf(stderr, "%s\n", bw_last_error());
    return 1;
  }
  /* was: wavefront_align(wf_aligner, pattern, plen, text, tlen); */
  const int status = bw_align(al, pattern, plen, text, tlen);
  if (status != BW_STATUS_ALG_COMPLETED) {
    fprintf(stderr, "status %d: %s\n", status, bw_strerror(status));
    bw_aligner_delete(al);
    return 1;
  }
  char cigar[8192];
  if (bw_cigar_sprint(al, cigar, (int)sizeof(cigar)) < 0) {
    fprintf(stderr, "CIGAR longer than the buffer\n");
    bw_aligner_delete(al);
    return 1;
  }
  printf("%d\t%s\n", bw_get_score(al), cigar); /* the .alg line */
  bw_aligner_delete(al);
  return 0;
}
