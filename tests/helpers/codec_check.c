/*
 * Test helper: typed reads of the first snapshot of a file through the
 * codec's fs_get_data (rigid and coerced), printed one line per check.
 *   codec_check FILE
 */
#include <stdio.h>
#include <stdlib.h>

#include "filestruct.h"

static const fs_item *first_set(fs_reader *r, fs_item *it)
{
    for (;;) {
        int st = fs_read_item(r, it);
        if (st <= 0) return NULL;
        if (fs_find(it, "Particles")) return it;
        fs_free_item(it);
    }
}

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    fs_reader r = {f, -1, ""};
    fs_item snap;
    if (!first_set(&r, &snap)) return 3;
    const fs_item *par = fs_find(fs_find(&snap, "Parameters"), "Nobj");
    const fs_item *pts = fs_find(&snap, "Particles");
    const fs_item *ps = fs_find(pts, "PhaseSpace"), *ms = fs_find(pts, "Mass");
    char err[256];
    int nobj = 0;
    if (fs_get_data(par, "i", &nobj, 0, NULL, 0, err, sizeof err)) return printf("nobj: %s\n", err), 4;
    const int dps[3] = {nobj, 2, 3}, dm[1] = {nobj}, dbad[3] = {nobj, 3, 2};
    double *pd = malloc(sizeof(double) * 6 * nobj), *md = malloc(sizeof(double) * nobj);
    int rc = fs_get_data(ps, "d", pd, 3, dps, 0, err, sizeof err);
    printf("rigid %d %s\n", rc, rc ? err : "ok");
    rc = fs_get_data(ps, "d", pd, 3, dps, 1, err, sizeof err);
    printf("coerced %d %s\n", rc, rc ? err : "ok");
    if (!rc) {
        printf("phase");
        for (int k = 0; k < 6 * nobj; k++) printf(" %.17g", pd[k]);
        printf("\n");
    }
    rc = fs_get_data(ms, "d", md, 1, dm, 1, err, sizeof err);
    printf("mass %d %s\n", rc, rc ? err : "ok");
    rc = fs_get_data(ps, "d", pd, 3, dbad, 1, err, sizeof err);
    printf("baddims %d %s\n", rc, rc ? err : "ok");
    rc = fs_get_data(ps, "d", pd, 0, NULL, 1, err, sizeof err);
    printf("scalar %d %s\n", rc, rc ? err : "ok");
    rc = fs_get_data(ps, "i", pd, 3, dps, 1, err, sizeof err);
    printf("toint %d %s\n", rc, rc ? err : "ok");
    fs_free_item(&snap);
    free(pd);
    free(md);
    fclose(f);
    return 0;
}
