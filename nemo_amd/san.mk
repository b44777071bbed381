# Host programs built with AddressSanitizer and UndefinedBehaviorSanitizer
# (gcc, CPU-side C only; they link the normal libnemo_bh.so).  Used by the CPU
# test tests/test_sanitizers.py:  make -C nemo_amd -f san.mk
CC       ?= gcc
SANFLAGS = -O1 -g -fno-omit-frame-pointer -fsanitize=address,undefined -fno-sanitize-recover=undefined
CFLAGS_CLI = -std=gnu11 -Wall -Wextra -Wno-unused-parameter
CLI      = cli/hackforce.c cli/hackcode1.c cli/csf.c cli/filestruct.c cli/filestruct.h cli/getparam.c cli/getparam.h

all: bin-san/hackforce bin-san/hackcode1 bin-san/csf bin-san/bhtree.so

bin-san:
	mkdir -p bin-san
bin-san/hackforce: $(CLI) ../include/bh.h | bin-san
	$(CC) $(CFLAGS_CLI) $(SANFLAGS) -o $@ cli/hackforce.c cli/filestruct.c cli/getparam.c -L. -lnemo_bh -Wl,-rpath,'$$ORIGIN/..'
bin-san/hackcode1: $(CLI) ../include/bh.h | bin-san
	$(CC) $(CFLAGS_CLI) $(SANFLAGS) -o $@ cli/hackcode1.c cli/filestruct.c cli/getparam.c -L. -lnemo_bh -lm -Wl,-rpath,'$$ORIGIN/..'
bin-san/csf: $(CLI) | bin-san
	$(CC) $(CFLAGS_CLI) $(SANFLAGS) -o $@ cli/csf.c cli/filestruct.c
bin-san/bhtree.so: cli/acc_bhtree.c ../include/bh.h | bin-san
	$(CC) $(CFLAGS_CLI) $(SANFLAGS) -shared -fPIC -o $@ cli/acc_bhtree.c -L. -lnemo_bh -Wl,-rpath,'$$ORIGIN/..'
