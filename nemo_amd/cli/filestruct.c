/*
 * filestruct.c -- reader and writer of NEMO structured binary files; the
 * item grammar is described in filestruct.h.
 */
#include "filestruct.h"

#include <stdlib.h>
#include <string.h>

#define FS_SINGLE 0x0992
#define FS_PLURAL 0x0B92

static int fail(fs_reader *r, const char *msg, const char *tag)
{
    if (tag && *tag) snprintf(r->err, sizeof r->err, "%s (tag %s)", msg, tag);
    else snprintf(r->err, sizeof r->err, "%s", msg);
    return -1;
}

static size_t elsize_of(const char *type)
{
    if (!strcmp(type, "c") || !strcmp(type, "b") || !strcmp(type, "a")) return 1;
    if (!strcmp(type, "s") || !strcmp(type, "h")) return 2;
    if (!strcmp(type, "i") || !strcmp(type, "f")) return 4;
    if (!strcmp(type, "l") || !strcmp(type, "d")) return 8;
    return 0;
}

static uint16_t bswap16(uint16_t v) { return (uint16_t)((v >> 8) | (v << 8)); }

static void swap_elems(void *p, size_t count, size_t es)
{
    unsigned char *b = (unsigned char *)p;
    if (es < 2) return;
    for (size_t i = 0; i < count; i++, b += es)
        for (size_t k = 0; k < es / 2; k++) {
            unsigned char t = b[k];
            b[k] = b[es - 1 - k];
            b[es - 1 - k] = t;
        }
}

/* NUL-terminated string of at most cap-1 characters. */
static int read_cstr(fs_reader *r, char *s, size_t cap, const char *what)
{
    for (size_t i = 0;; i++) {
        int c = fgetc(r->f);
        if (c == EOF) return fail(r, "truncated item header", what);
        if (i + 1 >= cap) return fail(r, "item name too long", what);
        s[i] = (char)c;
        if (c == 0) return 0;
    }
}

static int add_kid(fs_item *set, const fs_item *kid)
{
    if (set->nkids == set->capkids) {
        int nc = set->capkids ? 2 * set->capkids : 8;
        fs_item *k = (fs_item *)realloc(set->kids, (size_t)nc * sizeof *k);
        if (!k) return -1;
        set->kids = k;
        set->capkids = nc;
    }
    set->kids[set->nkids++] = *kid;
    return 0;
}

/* Returns 1 item, 2 end-of-set, 0 clean EOF (top level only), -1 error. */
static int read_one(fs_reader *r, fs_item *it, int depth)
{
    memset(it, 0, sizeof *it);
    unsigned char mb[2];
    size_t got = fread(mb, 1, 2, r->f);
    if (got == 0 && depth == 0 && feof(r->f)) return 0;
    if (got != 2) return fail(r, "truncated item header", NULL);
    uint16_t m;
    memcpy(&m, mb, 2);
    int swap;
    if (m == FS_SINGLE || m == FS_PLURAL) swap = 0;
    else if (bswap16(m) == FS_SINGLE || bswap16(m) == FS_PLURAL) swap = 1;
    else return fail(r, "not a structured file (bad magic number)", NULL);
    if (r->swap < 0) r->swap = swap;
    else if (r->swap != swap) return fail(r, "mixed byte order in one stream", NULL);
    if (swap) m = bswap16(m);
    if (read_cstr(r, it->type, sizeof it->type, "type")) return -1;
    if (!strcmp(it->type, ")")) return depth > 0 ? 2 : fail(r, "end of set without a set", NULL);
    if (read_cstr(r, it->tag, sizeof it->tag, it->type)) return -1;
    if (m == FS_PLURAL) {
        for (;;) {
            int32_t d;
            if (fread(&d, 4, 1, r->f) != 1) return fail(r, "truncated dimensions", it->tag);
            if (swap) swap_elems(&d, 1, 4);
            if (d == 0) break;
            if (d < 0) return fail(r, "negative dimension", it->tag);
            if (it->ndim == FS_MAXDIM) return fail(r, "too many dimensions", it->tag);
            it->dims[it->ndim++] = d;
        }
    }
    if (!strcmp(it->type, "(")) {
        if (depth > 64) return fail(r, "sets nested too deeply", it->tag);
        for (;;) {
            fs_item kid;
            int s = read_one(r, &kid, depth + 1);
            if (s < 0) {
                fs_free_item(&kid);
                return -1;
            }
            if (s == 2) break;
            if (add_kid(it, &kid)) {
                fs_free_item(&kid);
                return fail(r, "out of memory", it->tag);
            }
        }
        return 1;
    }
    it->elsize = elsize_of(it->type);
    if (!it->elsize) return fail(r, "unknown item type", it->tag);
    size_t count = 1;
    for (int k = 0; k < it->ndim; k++) {
        if (count > ((size_t)1 << 40) / (size_t)it->dims[k]) return fail(r, "item too large", it->tag);
        count *= (size_t)it->dims[k];
    }
    it->count = count;
    it->data = malloc(count * it->elsize + 1);
    if (!it->data) return fail(r, "out of memory", it->tag);
    if (count && fread(it->data, it->elsize, count, r->f) != count) return fail(r, "truncated item data", it->tag);
    ((char *)it->data)[count * it->elsize] = 0;
    if (swap) swap_elems(it->data, count, it->elsize);
    return 1;
}

int fs_read_item(fs_reader *r, fs_item *it)
{
    int s = read_one(r, it, 0);
    if (s < 0) fs_free_item(it);
    return s;
}

void fs_free_item(fs_item *it)
{
    for (int k = 0; k < it->nkids; k++) fs_free_item(&it->kids[k]);
    free(it->kids);
    free(it->data);
    it->kids = NULL;
    it->data = NULL;
    it->nkids = it->capkids = 0;
}

const fs_item *fs_find(const fs_item *set, const char *tag)
{
    for (int k = 0; k < set->nkids; k++)
        if (!strcmp(set->kids[k].tag, tag)) return &set->kids[k];
    return NULL;
}

static int append_hist(char **hist, size_t *len, int *n, const char *text)
{
    size_t l = strlen(text) + 1;
    char *h = (char *)realloc(*hist, *len + l);
    if (!h) return -1;
    memcpy(h + *len, text, l);
    *hist = h;
    *len += l;
    (*n)++;
    return 0;
}

/* Rigid types: hackforce's get_data accepts only the declared type
 * (filesecret.c:589-606). */
static int want(fs_reader *r, const fs_item *it, const char *tag, const char *type, size_t count)
{
    if (!it) return fail(r, "get_data: missing item", tag);
    if (strcmp(it->type, type)) {
        snprintf(r->err, sizeof r->err, "get_data: %s has type %s, %s expected", tag, it->type, type);
        return -1;
    }
    if (it->count != count) {
        snprintf(r->err, sizeof r->err, "get_data: %s holds %zu values, %zu expected", tag, it->count, count);
        return -1;
    }
    return 0;
}

int fs_read_snapshot(fs_reader *r, fs_snapshot *s, char **hist, size_t *histlen, int *nhist)
{
    memset(s, 0, sizeof *s);
    for (;;) {
        fs_item it;
        int st = fs_read_item(r, &it);
        if (st <= 0) return st;
        if ((!strcmp(it.tag, "History") || !strcmp(it.tag, "Headline")) && !strcmp(it.type, "c")) {
            if (hist && append_hist(hist, histlen, nhist, (const char *)it.data)) {
                fs_free_item(&it);
                return fail(r, "out of memory", "History");
            }
            fs_free_item(&it);
            continue;
        }
        if (strcmp(it.tag, "SnapShot") || strcmp(it.type, "(")) {
            fs_free_item(&it);
            continue;
        }
        const fs_item *par = fs_find(&it, "Parameters"), *prt = fs_find(&it, "Particles");
        if (!par || !prt) {
            fs_free_item(&it);
            continue;
        }
        int rc = -1;
        const fs_item *nobj = fs_find(par, "Nobj"), *tm = fs_find(par, "Time");
        const fs_item *cs = fs_find(prt, "CoordSystem");
        if (want(r, nobj, "Nobj", "i", 1)) goto done;
        int32_t n = *(const int32_t *)nobj->data;
        if (n < 1) {
            snprintf(r->err, sizeof r->err, "read_snapshot: Nobj = %d  is absurd", (int)n);
            goto done;
        }
        if (tm) {
            if (want(r, tm, "Time", "d", 1)) goto done;
            s->time = *(const double *)tm->data;
        }
        if (want(r, cs, "CoordSystem", "i", 1)) goto done;
        if (*(const int32_t *)cs->data != FS_CS_CARTESIAN_3D_PHASE) {
            snprintf(r->err, sizeof r->err, "read_snapshot: cannot handle CoordSystem = %d", *(const int32_t *)cs->data);
            goto done;
        }
        const fs_item *ms = fs_find(prt, "Mass"), *ps = fs_find(prt, "PhaseSpace");
        if (want(r, ms, "Mass", "d", (size_t)n) || want(r, ps, "PhaseSpace", "d", (size_t)n * 6)) goto done;
        s->n = n;
        s->mass = (double *)malloc((size_t)n * sizeof(double));
        s->pos = (double *)malloc((size_t)n * 3 * sizeof(double));
        s->vel = (double *)malloc((size_t)n * 3 * sizeof(double));
        if (!s->mass || !s->pos || !s->vel) {
            fail(r, "out of memory", "Particles");
            fs_free_snapshot(s);
            goto done;
        }
        memcpy(s->mass, ms->data, (size_t)n * sizeof(double));
        const double *p = (const double *)ps->data;
        for (int32_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                s->pos[3 * (size_t)i + k] = p[6 * (size_t)i + k];
                s->vel[3 * (size_t)i + k] = p[6 * (size_t)i + 3 + k];
            }
        rc = 1;
    done:
        fs_free_item(&it);
        return rc;
    }
}

void fs_free_snapshot(fs_snapshot *s)
{
    free(s->mass);
    free(s->pos);
    free(s->vel);
    s->mass = s->pos = s->vel = NULL;
}

static int put_head(FILE *f, uint16_t magic, const char *type, const char *tag)
{
    if (fwrite(&magic, 2, 1, f) != 1) return -1;
    if (fwrite(type, strlen(type) + 1, 1, f) != 1) return -1;
    if (tag && fwrite(tag, strlen(tag) + 1, 1, f) != 1) return -1;
    return 0;
}

int fs_put_set(FILE *f, const char *tag) { return put_head(f, FS_SINGLE, "(", tag); }

int fs_put_tes(FILE *f) { return put_head(f, FS_SINGLE, ")", NULL); }

int fs_put_data(FILE *f, const char *tag, const char *type, const void *data, int ndim, const int *dims)
{
    size_t es = elsize_of(type), count = 1;
    if (!es || ndim < 0 || ndim > FS_MAXDIM) return -1;
    if (put_head(f, ndim ? FS_PLURAL : FS_SINGLE, type, tag)) return -1;
    for (int k = 0; k < ndim; k++) {
        int32_t d = dims[k];
        if (fwrite(&d, 4, 1, f) != 1) return -1;
        count *= (size_t)dims[k];
    }
    if (ndim) {
        int32_t z = 0;
        if (fwrite(&z, 4, 1, f) != 1) return -1;
    }
    if (count && fwrite(data, es, count, f) != count) return -1;
    return 0;
}

int fs_put_history(FILE *f, const char *text)
{
    int d = (int)strlen(text) + 1;
    return fs_put_data(f, "History", "c", text, 1, &d);
}

int fs_put_item(FILE *f, const fs_item *it)
{
    if (!strcmp(it->type, "(")) {
        if (fs_put_set(f, it->tag)) return -1;
        for (int k = 0; k < it->nkids; k++)
            if (fs_put_item(f, &it->kids[k])) return -1;
        return fs_put_tes(f);
    }
    return fs_put_data(f, it->tag, it->type, it->data, it->ndim, it->dims);
}

/* ---- typed reads and conversions ------------------------------------------ */

int fs_get_data(const fs_item *it, const char *type, void *dat, int ndim, const int *dims, int coerce, char *err,
                size_t errlen)
{
    if (!it || !strcmp(it->type, "(")) {
        snprintf(err, errlen, "get_data: item %s: not a data item", it ? it->tag : "?");
        return -1;
    }
    const int same = !strcmp(it->type, type);
    const int f2d = !strcmp(it->type, "f") && !strcmp(type, "d");
    const int d2f = !strcmp(it->type, "d") && !strcmp(type, "f");
    if (!same && !(coerce && (f2d || d2f))) {
        snprintf(err, errlen, coerce ? "get_data_sub: item %s: types %s, %s don't convert"
                                     : "get_data_sub: item %s: types %s, %s don't match",
                 it->tag, it->type, type);
        return -1;
    }
    if (ndim > 0 && it->ndim > 0) {
        int ok = ndim == it->ndim;
        for (int k = 0; ok && k < ndim; k++) ok = dims[k] == it->dims[k];
        if (!ok) {
            snprintf(err, errlen, "get_data_sub: item %s: dimensions don't match", it->tag);
            return -1;
        }
    } else if (ndim == 0 && it->ndim > 0) {
        snprintf(err, errlen, "get_data_sub: item %s: can't copy plural to scalar", it->tag);
        return -1;
    } else if (ndim > 0 && it->ndim == 0) {
        snprintf(err, errlen, "get_data_sub: item %s: can't copy scalar to plural", it->tag);
        return -1;
    }
    if (same) {
        memcpy(dat, it->data, it->count * it->elsize);
    } else if (f2d) {
        const float *s = (const float *)it->data;
        double *d = (double *)dat;
        for (size_t i = 0; i < it->count; i++) d[i] = (double)s[i];
    } else {
        const double *s = (const double *)it->data;
        float *d = (float *)dat;
        for (size_t i = 0; i < it->count; i++) d[i] = (float)s[i];
    }
    return 0;
}

/* First code that applies to an item of type t (findtype). */
static const char *find_code(const char *const *codes, int ncodes, const char *t)
{
    for (int i = 0; i < ncodes; i++) {
        const char c = codes[i][0];
        if ((!strcmp(t, "d") && c == 'd') || (!strcmp(t, "f") && c == 'f') || (!strcmp(t, "h") && c == 'h') ||
            (!strcmp(t, "i") && (c == 'i' || c == 's')))
            return codes[i];
    }
    return NULL;
}

int fs_convert(fs_item *it, const char *const *codes, int ncodes, void (*warn)(const char *code, const char *tag))
{
    if (!strcmp(it->type, "(")) {
        int n = 0;
        for (int k = 0; k < it->nkids; k++) {
            const int c = fs_convert(&it->kids[k], codes, ncodes, warn);
            if (c < 0) return -1;
            n += c;
        }
        return n;
    }
    const char *cp = find_code(codes, ncodes, it->type);
    if (!cp) return 0;
    const int d2f = !strcmp(it->type, "d") && !strcmp(cp, "d2f");
    const int f2d = !strcmp(it->type, "f") && !strcmp(cp, "f2d");
    if (!d2f && !f2d) {
        if (warn) warn(cp, it->tag);
        return 0;
    }
    const size_t es = d2f ? sizeof(float) : sizeof(double);
    void *nd = malloc(it->count * es + 1);
    if (!nd) return -1;
    if (d2f) {
        const double *s = (const double *)it->data;
        for (size_t i = 0; i < it->count; i++) ((float *)nd)[i] = (float)s[i];
    } else {
        const float *s = (const float *)it->data;
        for (size_t i = 0; i < it->count; i++) ((double *)nd)[i] = (double)s[i];
    }
    free(it->data);
    it->data = nd;
    it->elsize = es;
    snprintf(it->type, sizeof it->type, "%s", d2f ? "f" : "d");
    return 1;
}
