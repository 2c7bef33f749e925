/*
 * inp.h -- reader for the reference's INI-like .inp format
 * (src/io/InputParser.c:225-478): "[section]" headers, "key = value" lines,
 * comment lines starting with one of  " # / - * , trailing comments cut at
 * one of  # * "  (plus / for EOS parameter files), list values "[a, b, c]"
 * whose k-th element is read by the *_k getters.  Keys are "section.key"
 * unless the key already contains a dot; lookup is case-insensitive.
 */
#ifndef SALEC_INP_H
#define SALEC_INP_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct inp_file inp_file;

inp_file *inp_open(const char *path, int eos_style);   /* NULL on error      */
void      inp_close(inp_file *f);
int       inp_has(const inp_file *f, const char *key);
double    inp_d(const inp_file *f, const char *key, double dflt);
int       inp_i(const inp_file *f, const char *key, int dflt);
/* copy string value (trimmed) into out[n]; returns 1 if found */
int       inp_s(const inp_file *f, const char *key, char *out, int n, const char *dflt);
/* k-th element of a "[a,b,...]" list; empty or missing elements give dflt */
double    inp_dk(const inp_file *f, const char *key, int k, double dflt);
int       inp_ik(const inp_file *f, const char *key, int k, int dflt);
int       inp_sk(const inp_file *f, const char *key, int k, char *out, int n, const char *dflt);

#ifdef __cplusplus
}
#endif
#endif
