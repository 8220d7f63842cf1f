/*
 * cfl.h -- BART's array files: "<name>.hdr" + "<name>.cfl" pairs and ".ra"
 * files (misc/io.c:380-468 header, :960-1040 ra; misc/mmio.c load/create).
 *
 * A .hdr file is text: "# Dimensions\n" followed by the dims separated by
 * spaces; further "# Keyword" sections are skipped.  A .cfl file holds the
 * raw complex float data, column-major.  A .ra file has a 6 x uint64 header
 * (magic 0x7961727261776172, flags, element type 4 = complex, element size
 * 8, data bytes, number of dims), the dims as uint64, then the data.
 */
#ifndef BAM_CLI_CFL_H
#define BAM_CLI_CFL_H

#include <complex.h>

/* reads the array into newly allocated memory; dims beyond the file's are 1;
 * a file with more than D non-trivial dims is an error */
complex float* cfl_load(const char* name, int D, long dims[]);

/* writes an array (header and data) */
void cfl_write(const char* name, int D, const long dims[], const complex float* data);

long cfl_size(int D, const long dims[]);

/* BART's error(): message to stderr, exit status 1 */
void cli_error(const char* fmt, ...) __attribute__((noreturn, format(printf, 1, 2)));

/* the command line written into .hdr files ("# Command") */
void cfl_set_command(int argc, char* argv[]);

#endif
