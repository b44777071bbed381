# The oracle built with AddressSanitizer and UndefinedBehaviorSanitizer into
# oracle/_san/liboracle.so (CPU checks only; loaded through NEMO_ORACLE_LIB by
# tests/test_sanitizers.py).  make -C oracle -f san.mk
CC      ?= gcc
CFLAGS  = -O1 -g -fno-omit-frame-pointer -fPIC -ffp-contract=off -fno-fast-math -std=gnu11 \
          -fsanitize=address,undefined -fno-sanitize-recover=undefined

_san/liboracle.so: bh_oracle.c tc_oracle.c lc_oracle.c le_oracle.c
	mkdir -p _san
	$(CC) $(CFLAGS) -shared -o $@ bh_oracle.c tc_oracle.c lc_oracle.c le_oracle.c -lm -lpthread
