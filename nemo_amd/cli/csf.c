/*
 * csf -- copy a structured file item by item (NEMO's csf, src/kernel/io/
 * csf.c, with in=, out= and convert=): a byte-swapped input is written back
 * in host byte order; convert= takes a comma- or space-separated list of
 * conversion codes, of which d2f and f2d are carried out (copy_item_cvt,
 * filesecret.c:127-240; other codes are reported with a warning and the item
 * is copied as it is).  The host-side check of the snapshot codec that
 * hackforce uses.  Errors print "### Fatal error [csf]: ..." and exit with
 * status 255.
 */
#include <stdio.h>
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

int main(int argc, char **argv)
{
    const char *in = NULL, *out = NULL;
    char cvtbuf[256] = "";
    const char *codes[16];
    int ncodes = 0, npos = 0;
    for (int i = 1; i < argc; i++) {
        if (!strncmp(argv[i], "in=", 3)) in = argv[i] + 3;
        else if (!strncmp(argv[i], "out=", 4)) out = argv[i] + 4;
        else if (!strncmp(argv[i], "convert=", 8)) snprintf(cvtbuf, sizeof cvtbuf, "%s", argv[i] + 8);
        else if (!strchr(argv[i], '=') && npos == 0) in = argv[i], npos++;
        else if (!strchr(argv[i], '=') && npos == 1) out = argv[i], npos++;
        else return fatal("unknown argument ", argv[i]);
    }
    if (!in || !out) return fatal("usage: csf in=FILE out=FILE [convert=d2f,f2d]", NULL);
    for (char *t = strtok(cvtbuf, ", "); t; t = strtok(NULL, ", ")) {
        if (ncodes == (int)(sizeof codes / sizeof codes[0])) return fatal("convert=: too many codes", NULL);
        codes[ncodes++] = t;
    }
    struct stat sb;
    if (strcmp(out, "-") && stat(out, &sb) == 0) return fatal("stropen: file already exists: ", out);
    FILE *fi = strcmp(in, "-") ? fopen(in, "rb") : stdin;
    if (!fi) return fatal("stropen: cannot open ", in);
    FILE *fo = strcmp(out, "-") ? fopen(out, "wb") : stdout;
    if (!fo) return fatal("stropen: cannot open ", out);
    fs_reader r = {fi, -1, ""};
    for (;;) {
        fs_item it;
        int st = fs_read_item(&r, &it);
        if (st < 0) return fatal(r.err, NULL);
        if (st == 0) break;
        if (ncodes && fs_convert(&it, codes, ncodes, warn_cvt) < 0) return fatal("convert=: out of memory", NULL);
        if (fs_put_item(fo, &it)) return fatal("write error on ", out);
        fs_free_item(&it);
    }
    if (fo != stdout && fclose(fo)) return fatal("write error on ", out);
    return 0;
}
