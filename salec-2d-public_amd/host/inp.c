/*
 * inp.c -- .inp reader (format of src/io/InputParser.c:225-478).
 */
#include "inp.h"
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>

typedef struct { char key[160]; char val[512]; } inp_kv;
struct inp_file { int n, cap; inp_kv *kv; };

static void trim_in_place(char *s)
{
    size_t n = strlen(s);
    while (n > 0 && isspace((unsigned char)s[n - 1])) s[--n] = '\0';
    size_t i = 0;
    /* leading whitespace and dots are dropped, as the reference's ltrim does */
    while (s[i] && (isspace((unsigned char)s[i]) || s[i] == '.')) i++;
    if (i) memmove(s, s + i, strlen(s + i) + 1);
}

static int is_comment_line(const char *s)
{
    return s[0] == '\0' || strchr("\"#/-*", s[0]) != NULL;
}

inp_file *inp_open(const char *path, int eos_style)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return NULL;
    inp_file *f = (inp_file *)calloc(1, sizeof(inp_file));
    if (!f) { fclose(fp); return NULL; }
    char section[160] = "mesh";
    char line[1024];
    const char *cut = eos_style ? "/#*\"" : "#*\"";
    while (fgets(line, sizeof line, fp)) {
        line[strcspn(line, "\r\n")] = '\0';
        trim_in_place(line);
        if (is_comment_line(line)) continue;
        size_t L = strlen(line);
        if (line[0] == '[' && line[L - 1] == ']') {
            line[L - 1] = '\0';
            snprintf(section, sizeof section, "%s", line + 1);
            continue;
        }
        /* key: first token delimited by '=' or blanks */
        char key[160] = {0}, val[512] = {0};
        const char *p = line;
        while (*p && (*p == '=' || *p == ' ')) p++;
        int k = 0;
        while (*p && *p != '=' && *p != ' ' && k < (int)sizeof(key) - 1) key[k++] = *p++;
        while (*p && (*p == '=' || *p == ' ')) p++;
        int v = 0;
        while (*p && !strchr(cut, *p) && v < (int)sizeof(val) - 1) val[v++] = *p++;
        trim_in_place(key);
        trim_in_place(val);
        if (!key[0] || !val[0]) continue;
        if (f->n == f->cap) {
            int nc = f->cap ? 2 * f->cap : 64;
            inp_kv *nk = (inp_kv *)realloc(f->kv, sizeof(inp_kv) * nc);
            if (!nk) break;
            f->kv = nk; f->cap = nc;
        }
        if (strchr(key, '.')) snprintf(f->kv[f->n].key, sizeof f->kv[f->n].key, "%s", key);
        else snprintf(f->kv[f->n].key, sizeof f->kv[f->n].key, "%s.%s", section, key);
        snprintf(f->kv[f->n].val, sizeof f->kv[f->n].val, "%s", val);
        f->n++;
    }
    fclose(fp);
    return f;
}

void inp_close(inp_file *f)
{
    if (!f) return;
    free(f->kv);
    free(f);
}

static const char *lookup(const inp_file *f, const char *key)
{
    const char *r = NULL;
    for (int i = 0; i < f->n; ++i)
        if (strcasecmp(f->kv[i].key, key) == 0) r = f->kv[i].val;
    return r;
}

int inp_has(const inp_file *f, const char *key) { return lookup(f, key) != NULL; }

double inp_d(const inp_file *f, const char *key, double dflt)
{
    const char *v = lookup(f, key);
    return v ? strtod(v, NULL) : dflt;
}

int inp_i(const inp_file *f, const char *key, int dflt)
{
    const char *v = lookup(f, key);
    return v ? (int)strtol(v, NULL, 10) : dflt;
}

int inp_s(const inp_file *f, const char *key, char *out, int n, const char *dflt)
{
    const char *v = lookup(f, key);
    snprintf(out, n, "%s", v ? v : dflt);
    return v != NULL;
}

/* k-th comma-separated element of "[...]"; returns 0 if absent/empty */
static int list_elem(const char *v, int k, char *out, int n)
{
    size_t L = strlen(v);
    out[0] = '\0';
    if (L < 2 || v[0] != '[' || v[L - 1] != ']') return 0;
    int idx = 0, o = 0;
    for (size_t i = 1; i + 1 < L; ++i) {
        if (v[i] == ',') { idx++; if (idx > k) break; continue; }
        if (idx == k && o < n - 1) out[o++] = v[i];
    }
    out[o] = '\0';
    trim_in_place(out);
    return out[0] != '\0';
}

double inp_dk(const inp_file *f, const char *key, int k, double dflt)
{
    const char *v = lookup(f, key);
    char tmp[256];
    if (!v || !list_elem(v, k, tmp, sizeof tmp)) return dflt;
    return strtod(tmp, NULL);
}

int inp_ik(const inp_file *f, const char *key, int k, int dflt)
{
    const char *v = lookup(f, key);
    char tmp[256];
    if (!v || !list_elem(v, k, tmp, sizeof tmp)) return dflt;
    return (int)strtol(tmp, NULL, 10);
}

int inp_sk(const inp_file *f, const char *key, int k, char *out, int n, const char *dflt)
{
    const char *v = lookup(f, key);
    char tmp[256];
    if (!v || !list_elem(v, k, tmp, sizeof tmp)) { snprintf(out, n, "%s", dflt); return 0; }
    snprintf(out, n, "%s", tmp);
    return 1;
}
