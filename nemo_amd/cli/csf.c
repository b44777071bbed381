/*
 * csf -- copy a structured file item by item (NEMO's csf, in= out= only):
 * a byte-swapped input is written back in host byte order.  The host-side
 * check of the snapshot codec that hackforce uses.  Errors print
 * "### Fatal error [csf]: ..." and exit with status 255.
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

int main(int argc, char **argv)
{
    const char *in = NULL, *out = NULL;
    int npos = 0;
    for (int i = 1; i < argc; i++) {
        if (!strncmp(argv[i], "in=", 3)) in = argv[i] + 3;
        else if (!strncmp(argv[i], "out=", 4)) out = argv[i] + 4;
        else if (!strchr(argv[i], '=') && npos == 0) in = argv[i], npos++;
        else if (!strchr(argv[i], '=') && npos == 1) out = argv[i], npos++;
        else return fatal("unknown argument ", argv[i]);
    }
    if (!in || !out) return fatal("usage: csf in=FILE out=FILE", NULL);
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
        if (fs_put_item(fo, &it)) return fatal("write error on ", out);
        fs_free_item(&it);
    }
    if (fo != stdout && fclose(fo)) return fatal("write error on ", out);
    return 0;
}
