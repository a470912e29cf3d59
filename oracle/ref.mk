# Builds oracle/_ref/libluxref.so from the reference's own host sources, in
# place under $(REF_SRC), plus the extern "C" shim oracle/ref_driver.cpp.
# Flags follow the reference's default Release configuration (C++20, -O3).
# Usage (from the repository root): make -f oracle/ref.mk
REF_SRC ?= /root/reference/src
OUT := oracle/_ref

# The reference translation units on the CPU path-tracer path.
REF_TUS := $(REF_SRC)/integrator/cpu_path_tracer.cpp $(REF_SRC)/light/light_distribution.cpp \
           $(REF_SRC)/scene/prepared_scene_cpu.cpp

$(OUT)/libluxref.so: oracle/ref_driver.cpp $(REF_TUS)
	mkdir -p $(OUT)
	g++ -std=c++20 -O3 -DNDEBUG -fPIC -shared -DLUX_ENABLE_CUDA=0 \
	    -DLUX_ENABLE_SCENE_DEEP_CURRENT_CHECKS=0 -I$(REF_SRC) \
	    oracle/ref_driver.cpp $(REF_TUS) -o $@

.PHONY: clean
clean:
	rm -f $(OUT)/libluxref.so
