# Host-only sanitizer build of the oracle's C restatement (AddressSanitizer +
# UndefinedBehaviorSanitizer, gcc): liboracle_san.so, loaded by
# tests/test_sanitized_host.py in a child process with libasan preloaded.
CC ?= gcc
SAN_CFLAGS = -O1 -g -std=c99 -fPIC -ffp-contract=off -fno-fast-math -fno-omit-frame-pointer \
             -fsanitize=address,undefined -fno-sanitize-recover=undefined -Wall -Wno-unused-function

liboracle_san.so: ntx_oracle.c ntx_oracle.h
	$(CC) $(SAN_CFLAGS) -shared -o $@ ntx_oracle.c -lm -lpthread
