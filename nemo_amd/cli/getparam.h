/*
 * getparam.h -- NEMO-style keyword handling for the host programs:
 * key=value arguments, positional values in keyword order before the first
 * key=value, "help" lists the keywords, fatal errors print
 * "### Fatal error [prog]: ..." and exit with status 255 (the behaviour of
 * NEMO's getparam.c and error.c:100-103,171-180).
 */
#ifndef NEMO_GETPARAM_H
#define NEMO_GETPARAM_H

typedef struct {
    const char *key, *def, *help;   /* def == NULL: the value must be given */
    const char *val;
} gp_keyword;

void gp_init(const char *prog, gp_keyword *keys, int nkeys, int argc, char **argv);
const char *gp_param(const char *key);
double gp_dparam(const char *key);
int gp_iparam(const char *key);
int gp_bparam(const char *key, int dflt);   /* blank -> dflt */
int gp_given(const char *key);              /* set on the command line */
int gp_choice(const char *key, const char *const *names, int n);   /* index of the value in names */
int gp_scanopt(const char *list, const char *word);
const char *gp_cmdline(void);                /* the invocation, for History */
void gp_fatal(const char *fmt, ...) __attribute__((format(printf, 1, 2), noreturn));

#endif
