/*
 * tools.h -- command-line tools of the "bart" drop-in (src/main.c tool
 * table counterparts): each takes argv with argv[0] = tool name.
 */
#ifndef BAM_CLI_TOOLS_H
#define BAM_CLI_TOOLS_H

#include <stdbool.h>

#define DIMS 16

/* dims of BART's 16-dim layout (misc/mri.h:9-27) */
enum { READ_DIM, PHS1_DIM, PHS2_DIM, COIL_DIM, MAPS_DIM, TE_DIM, COEFF_DIM };

int main_ecalib(int argc, char* argv[]);
int main_nufft(int argc, char* argv[]);
int main_pics(int argc, char* argv[]);
int main_traj(int argc, char* argv[]);

/* shared helpers (cli_util.c) */
int debug_level(void);
void set_debug_level(int l);
void debug_printf(int level, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
double timestamp(void);
void parse_vec3(const char* s, long v[3], const char* opt);
float parse_float(const char* s, const char* opt);
int parse_int(const char* s, const char* opt);
void unsupported_option(const char* tool, const char* opt);

/* --nufft-conf sub-options (noncart/nufft.c:100-113, parsed as
 * misc/opts.c:1227-1300 does: comma-separated short or long names, with
 * =value for the two that take one) */
struct nufft_conf_s;
void parse_nufft_conf(const char* arg, struct nufft_conf_s* conf);

/* misc/mri2.c:231-264 */
void estimate_im_dims(const long traj_dims[DIMS], const _Complex float* traj, long dims[3]);

/* misc/mri2.c:50-69: 1 where any coil is non-zero, else 0 */
void estimate_pattern(const long ksp_dims[DIMS], _Complex float* pattern, const _Complex float* ksp);

/* sense/optcom.c:74-109 (rescale = 1, compat = false, p = -1); sorts x */
float estimate_scaling_norm(long n, _Complex float* x);

#endif
