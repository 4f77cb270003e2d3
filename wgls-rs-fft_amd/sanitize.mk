# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build of the
# runtime (every csrc/*.hip; the sanitizers instrument host code only, the
# gfx950 device code is built as in build.py) and of the C++ mirror tests,
# into wgfft/_lib/san/.  Used by tests/test_sanitizers.py:
#   make -f wgls-rs-fft_amd/sanitize.mk -j8
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
HIPCC ?= /opt/rocm/bin/hipcc
CLANGXX ?= /opt/rocm/llvm/bin/clang++
OBJDIR := $(HERE)_build_san
OUT := $(HERE)wgfft/_lib/san
SRCS := $(wildcard $(HERE)csrc/*.hip)
OBJS := $(patsubst $(HERE)csrc/%.hip,$(OBJDIR)/%.o,$(SRCS))
HIPFLAGS := -O3 -fPIC -std=c++17 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
            -Wno-unused-command-line-argument -I $(HERE)../include

all: $(OUT)/libwgfft.so $(OUT)/gpu_fft_tests

$(OBJDIR)/%.o: $(HERE)csrc/%.hip $(wildcard $(HERE)csrc/*.hpp) $(HERE)../include/wgfft.h
	@mkdir -p $(OBJDIR)
	$(HIPCC) $(HIPFLAGS) -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=all -Xarch_host -fno-omit-frame-pointer -Xarch_host -g -c $< -o $@

$(OUT)/libwgfft.so: $(OBJS)
	@mkdir -p $(OUT)
	$(HIPCC) --offload-arch=gfx950 -shared -fPIC -pthread -Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -o $@ $(OBJS)

$(OUT)/gpu_fft_tests: $(HERE)cpp/gpu_fft_tests.cpp $(HERE)cpp/gpu_fft.hpp $(OUT)/libwgfft.so
	$(CLANGXX) -O1 -g -std=c++17 -fsanitize=address,undefined -fno-sanitize-recover=all -fno-omit-frame-pointer \
	    -o $@ $< -L$(OUT) -lwgfft -Wl,-rpath,'$$ORIGIN'

clean:
	rm -rf $(OBJDIR) $(OUT)

.PHONY: all clean
