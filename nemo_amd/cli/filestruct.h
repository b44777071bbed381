/*
 * filestruct.h -- NEMO structured binary files (snapshots) for the host
 * programs over the C ABI.
 *
 * Item stream (src/kernel/io/filesecret.c:912-966 writes, 1028-1137 reads;
 * inc/filestruct.h:15-49): a 2-byte magic (0x0992 singular item, 0x0B92
 * plural item), the NUL-terminated type name, the NUL-terminated tag (absent
 * for the end-of-set item ")"), for plural items a 0-terminated int vector of
 * dimensions, then the data.  A file whose magic reads byte-swapped is read
 * with every multi-byte value swapped (filesecret.c CHKSWAP).  Output is
 * written in host byte order, as the reference does.
 *
 * Snapshot layout (inc/snapshot/snapshot.h:22-56): SnapShot{ Parameters{Nobj,
 * Time}, Particles{CoordSystem, Mass, PhaseSpace, ...}, Diagnostics{...} },
 * preceded by History items.
 */
#ifndef NEMO_FILESTRUCT_H
#define NEMO_FILESTRUCT_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_MAXDIM 8
#define FS_CS_CARTESIAN_3D_PHASE 66306   /* CSCode(Cartesian, 3, 2) */

typedef struct fs_item {
    char type[8];                 /* "(" set, "c" "b" "s" "i" "l" "h" "f" "d" "a" */
    char tag[128];
    int ndim;                     /* 0: singular */
    int dims[FS_MAXDIM];
    size_t count;                 /* number of elements (1 for singular) */
    size_t elsize;
    void *data;                   /* host byte order; NULL for sets */
    struct fs_item *kids;         /* sets: children */
    int nkids, capkids;
} fs_item;

typedef struct fs_reader {
    FILE *f;
    int swap;                     /* -1 unknown (before the first item), 0/1 */
    char err[256];
} fs_reader;

/* One top-level item (a whole set with its children).  Returns 1, 0 at a
 * clean end of file, -1 on a malformed or truncated stream (r->err). */
int fs_read_item(fs_reader *r, fs_item *it);
void fs_free_item(fs_item *it);
const fs_item *fs_find(const fs_item *set, const char *tag);

/* A snapshot as hackforce reads it (hackforce.c:111-169). */
typedef struct fs_snapshot {
    int64_t n;
    double time;                  /* 0 when the set has no Time */
    double *mass;                 /* [n] */
    double *pos, *vel;            /* [n][3] */
} fs_snapshot;

/* Next SnapShot set holding Parameters and Particles; sets without them are
 * skipped (hackforce.c:121-143).  History/Headline text met on the way is
 * appended to *hist (a NUL-separated list; *nhist counts the strings).
 * Returns 1, 0 at end of stream, -1 on error (r->err). */
int fs_read_snapshot(fs_reader *r, fs_snapshot *s, char **hist, size_t *histlen, int *nhist);
void fs_free_snapshot(fs_snapshot *s);

/* Writers (host byte order). */
int fs_put_set(FILE *f, const char *tag);
int fs_put_tes(FILE *f);
int fs_put_data(FILE *f, const char *tag, const char *type, const void *data, int ndim, const int *dims);
int fs_put_history(FILE *f, const char *text);
/* Re-emit one parsed item (and its children) unchanged. */
int fs_put_item(FILE *f, const fs_item *it);

/* Typed read of one item (get_data / get_data_coerced, filesecret.c:532-608):
 * the stored type must equal `type`, or, with coerce, float and double
 * convert into each other (copyfun, filesecret.c:1204-1213; no other pair).
 * The shape must equal dims[0..ndim) (ndim 0: a singular item); plural to
 * scalar and scalar to plural are errors.  Returns 0, or -1 with the reason
 * in err. */
int fs_get_data(const fs_item *it, const char *type, void *dat, int ndim, const int *dims, int coerce, char *err,
                size_t errlen);

/* csf's convert= (copy_item_cvt, filesecret.c:127-240): every data item of
 * `it` (sets are recursed) whose type matches the first letter of one of the
 * codes (findtype, filesecret.c:1454-1472: d, f, h; i for codes starting
 * with i or s) is converted by the first such code: d2f (double to float),
 * f2d (float to double).  Any other matching code leaves the item as it is
 * and is reported through warn(code, tag).  Returns the number of items
 * converted, or -1 when memory runs out. */
int fs_convert(fs_item *it, const char *const *codes, int ncodes, void (*warn)(const char *code, const char *tag));

#ifdef __cplusplus
}
#endif
#endif
