/*
 * csf -- copy a structured file item by item, with NEMO csf's keywords
 * (src/kernel/io/csf.c:27-34): in=, out=, item= (top-level tags to copy,
 * exact names; all by default), select= (sorted 1-based positions among the
 * matching items; all by default), convert= (comma- or space-separated codes
 * of which d2f and f2d are carried out as copy_item_cvt does,
 * filesecret.c:127-240; other codes are reported with a warning and the item
 * is copied as it is), blocksize= (raw copy in blocks of that many bytes)
 * and headline= (a Headline item written first).  A byte-swapped input is
 * written back in host byte order.  Like the reference it reports "On
 * top-level: R items  read, W items written" on stderr.  Errors print
 * "### Fatal error [csf]: ..." and exit with status 255.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "filestruct.h"

static int fatal(const char *msg, const char *arg)
{
    fprintf(stderr, "### Fatal error [csf]: %s%s\n", msg, arg ? arg : "");
    return 255;
}

static void warn_cvt(const char *code, const char *tag)
{
    fprintf(stderr, "### Warning [csf]: Cannot convert %s yet in %s\n", code, tag);
}

#define MAXSEL 512
#define MAXITEMS 64

int main(int argc, char **argv)
{
    const char *in = NULL, *out = NULL, *headline = NULL;
    char cvtbuf[256] = "", itembuf[1024] = "", selbuf[4096] = "";
    const char *codes[16], *items[MAXITEMS];
    int select[MAXSEL];
    int ncodes = 0, nitems = 0, nsel = 0, npos = 0;
    long blocksize = 0;
    for (int i = 1; i < argc; i++) {
        if (!strncmp(argv[i], "in=", 3)) in = argv[i] + 3;
        else if (!strncmp(argv[i], "out=", 4)) out = argv[i] + 4;
        else if (!strncmp(argv[i], "convert=", 8)) snprintf(cvtbuf, sizeof cvtbuf, "%s", argv[i] + 8);
        else if (!strncmp(argv[i], "item=", 5)) snprintf(itembuf, sizeof itembuf, "%s", argv[i] + 5);
        else if (!strncmp(argv[i], "select=", 7)) snprintf(selbuf, sizeof selbuf, "%s", argv[i] + 7);
        else if (!strncmp(argv[i], "headline=", 9)) headline = argv[i] + 9;
        else if (!strncmp(argv[i], "blocksize=", 10)) blocksize = atol(argv[i] + 10);
        else if (!strchr(argv[i], '=') && npos == 0) in = argv[i], npos++;
        else if (!strchr(argv[i], '=') && npos == 1) out = argv[i], npos++;
        else return fatal("unknown argument ", argv[i]);
    }
    if (!in || !out)
        return fatal("usage: csf in=FILE out=FILE [item=] [select=] [convert=d2f,f2d] [blocksize=] [headline=]", NULL);
    for (char *t = strtok(cvtbuf, ", "); t; t = strtok(NULL, ", ")) {
        if (ncodes == (int)(sizeof codes / sizeof codes[0])) return fatal("convert=: too many codes", NULL);
        codes[ncodes++] = t;
    }
    for (char *t = strtok(itembuf, ", "); t; t = strtok(NULL, ", ")) {
        if (nitems == MAXITEMS) return fatal("item=: too many items", NULL);
        items[nitems++] = t;
    }
    for (char *t = strtok(selbuf, ", "); t; t = strtok(NULL, ", ")) {
        if (nsel == MAXSEL) return fatal("select=: too many numbers", NULL);
        select[nsel] = atoi(t);
        if (select[nsel] < 1) return fatal("bad number for select=, must be > 0: ", t);
        if (nsel > 0 && select[nsel] < select[nsel - 1]) return fatal("bad number for select=, array must be sorted: ", t);
        nsel++;
    }
    struct stat sb;
    if (strcmp(out, "-") && stat(out, &sb) == 0) return fatal("stropen: file already exists: ", out);
    FILE *fi = strcmp(in, "-") ? fopen(in, "rb") : stdin;
    if (!fi) return fatal("stropen: cannot open ", in);
    FILE *fo = strcmp(out, "-") ? fopen(out, "wb") : stdout;
    if (!fo) return fatal("stropen: cannot open ", out);
    if (blocksize > 0) {   /* raw copy, no structure */
        char *buf = (char *)malloc((size_t)blocksize);
        if (!buf) return fatal("blocksize=: out of memory", NULL);
        size_t k;
        while ((k = fread(buf, 1, (size_t)blocksize, fi)) > 0)
            if (fwrite(buf, 1, k, fo) != k) return fatal("write error on ", out);
        free(buf);
        if (fo != stdout && fclose(fo)) return fatal("write error on ", out);
        return 0;
    }
    if (headline) {
        int d = (int)strlen(headline) + 1;
        if (fs_put_data(fo, "Headline", "c", headline, 1, &d)) return fatal("write error on ", out);
    }
    fs_reader r = {fi, -1, ""};
    int nread = 0, nwritten = 0;
    for (;;) {
        fs_item it;
        int st = fs_read_item(&r, &it);
        if (st < 0) return fatal(r.err, NULL);
        if (st == 0) break;
        int match = nitems == 0;
        for (int k = 0; k < nitems && !match; k++) match = !strcmp(items[k], it.tag);
        if (match) {
            nread++;
            if (nsel == 0 || select[nwritten] == nread) {
                if (ncodes && fs_convert(&it, codes, ncodes, warn_cvt) < 0)
                    return fatal("convert=: out of memory", NULL);
                if (fs_put_item(fo, &it)) return fatal("write error on ", out);
                nwritten++;
            }
        }
        fs_free_item(&it);
        if (nsel > 0 && nwritten >= nsel) break;   /* every selected item written */
    }
    fprintf(stderr, "On top-level: %d items  read, %d items written\n", nread, nwritten);
    if (fo != stdout && fclose(fo)) return fatal("write error on ", out);
    return 0;
}
