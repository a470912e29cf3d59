# Builds oracle/_ref/libluxref.so from the reference's own host sources, in
# place under $(REF_SRC), plus the extern "C" shim oracle/ref_driver.cpp.
# Flags follow the reference's default Release configuration (C++20, -O3).
# The OBJ loader's translation unit is compiled as C++17: the fast_float code
# of its bundled tinyobjloader trips a constexpr check of this image's g++ 11
# in C++20 mode, and nothing in that file depends on C++20.
# Usage (from the repository root): make -f oracle/ref.mk
REF_SRC ?= /root/reference/src
OUT := oracle/_ref
FLAGS := -O3 -DNDEBUG -fPIC -DLUX_ENABLE_CUDA=0 -DLUX_ENABLE_SCENE_DEEP_CURRENT_CHECKS=0 -I$(REF_SRC)

# The reference translation units on the CPU path-tracer path.
REF_TUS := $(REF_SRC)/integrator/cpu_path_tracer.cpp $(REF_SRC)/light/light_distribution.cpp \
           $(REF_SRC)/scene/prepared_scene_cpu.cpp

$(OUT)/libluxref.so: oracle/ref_driver.cpp $(REF_TUS) $(OUT)/obj_loader.o
	mkdir -p $(OUT)
	g++ -std=c++20 $(FLAGS) -shared oracle/ref_driver.cpp $(REF_TUS) $(OUT)/obj_loader.o -o $@

# The reference's OBJ loader (with the tinyobjloader source it bundles).
$(OUT)/obj_loader.o: $(REF_SRC)/util/obj_loader.cpp
	mkdir -p $(OUT)
	g++ -std=c++17 $(FLAGS) -c $(REF_SRC)/util/obj_loader.cpp -o $@

.PHONY: clean
clean:
	rm -f $(OUT)/libluxref.so $(OUT)/obj_loader.o
