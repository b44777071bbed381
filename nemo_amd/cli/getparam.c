/*
 * getparam.c -- keyword handling shared by the host programs (getparam.h).
 */
#include "getparam.h"

#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char *g_prog = "nemo";
static gp_keyword *g_keys;
static int g_nkeys;
static char *g_cmd;

void gp_fatal(const char *fmt, ...)
{
    va_list ap;
    fflush(stdout);
    fprintf(stderr, "### Fatal error [%s]: ", g_prog);
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fprintf(stderr, "\n");
    exit(255);
}

static gp_keyword *find(const char *k, size_t len)
{
    for (int i = 0; i < g_nkeys; i++)
        if (strlen(g_keys[i].key) == len && !strncmp(g_keys[i].key, k, len)) return &g_keys[i];
    return NULL;
}

void gp_init(const char *prog, gp_keyword *keys, int nkeys, int argc, char **argv)
{
    g_prog = prog;
    g_keys = keys;
    g_nkeys = nkeys;
    size_t len = 1;
    for (int i = 0; i < argc; i++) len += strlen(argv[i]) + 1;
    g_cmd = (char *)calloc(len, 1);
    const char *base = strrchr(argv[0], '/');
    strcpy(g_cmd, base ? base + 1 : argv[0]);
    int npos = 0, seen_kw = 0;
    for (int i = 1; i < argc; i++) {
        strcat(g_cmd, " ");
        strcat(g_cmd, argv[i]);
        if (!strcmp(argv[i], "help") || !strncmp(argv[i], "help=", 5) || !strcmp(argv[i], "--help")) {
            for (int k = 0; k < nkeys; k++)
                printf("%s=%s\n    %s\n", keys[k].key, keys[k].def ? keys[k].def : "???", keys[k].help);
            exit(0);
        }
        const char *eq = strchr(argv[i], '=');
        if (!eq) {
            if (seen_kw || npos >= nkeys) gp_fatal("Positional argument \"%s\" after keywords", argv[i]);
            keys[npos++].val = argv[i];
            continue;
        }
        seen_kw = 1;
        gp_keyword *w = find(argv[i], (size_t)(eq - argv[i]));
        if (!w) {
            /* NEMO system keywords that do not change the result */
            if (!strncmp(argv[i], "debug=", 6) || !strncmp(argv[i], "error=", 6) || !strncmp(argv[i], "yapp=", 5))
                continue;
            gp_fatal("Parameter \"%s\" unknown", argv[i]);
        }
        w->val = eq + 1;
    }
}

const char *gp_cmdline(void) { return g_cmd ? g_cmd : ""; }

int gp_given(const char *key)
{
    gp_keyword *w = find(key, strlen(key));
    return w && w->val != NULL;
}

const char *gp_param(const char *key)
{
    gp_keyword *w = find(key, strlen(key));
    if (!w) gp_fatal("no keyword %s", key);
    if (w->val) return w->val;
    if (!w->def) gp_fatal("Parameter \"%s\" must be given a value", key);
    return w->def;
}

double gp_dparam(const char *key)
{
    const char *s = gp_param(key);
    char *end;
    errno = 0;
    double v = strtod(s, &end);
    while (*end == ' ') end++;
    if (*s == 0 || *end || errno) gp_fatal("Parameter %s=%s is not a number", key, s);
    return v;
}

int gp_iparam(const char *key)
{
    const char *s = gp_param(key);
    char *end;
    errno = 0;
    long v = strtol(s, &end, 10);
    while (*end == ' ') end++;
    if (*s == 0 || *end || errno) gp_fatal("Parameter %s=%s is not an integer", key, s);
    return (int)v;
}

int gp_choice(const char *key, const char *const *names, int n)
{
    const char *s = gp_param(key);
    for (int i = 0; i < n; i++)
        if (!strcmp(s, names[i])) return i;
    gp_fatal("Parameter %s=%s is not one of the allowed values", key, s);
}

int gp_bparam(const char *key, int dflt)
{
    const char *s = gp_param(key);
    if (!*s) return dflt;
    if (strchr("tTyY1", s[0])) return 1;
    if (strchr("fFnN0", s[0])) return 0;
    gp_fatal("Parameter %s=%s is not a boolean", key, s);
}

int gp_scanopt(const char *list, const char *word)
{
    size_t n = strlen(word);
    const char *p = list;
    while (*p) {
        while (*p == ',' || *p == ' ') p++;
        const char *q = p;
        while (*q && *q != ',' && *q != ' ') q++;
        if ((size_t)(q - p) == n && !strncmp(p, word, n)) return 1;
        p = q;
    }
    return 0;
}
