# Builds the gfx950 library cufhedb_amd/libcufhedb_amd.so (HIP kernels +
# extern "C" boundary + host key generation) and the checker libraries under
# oracle/.  `make -j8` from the repository root; no cmake needed.

HIPCC ?= /opt/rocm/bin/hipcc
CXX ?= g++
ARCH ?= gfx950
HIPFLAGS := --offload-arch=$(ARCH) -O3 -std=c++17 -fPIC -Wall \
            -Xclang -target-feature -Xclang -packed-fp32-ops
CXXFLAGS := -O3 -std=c++17 -fPIC -Wall -Wextra

SRC := cufhedb_amd/csrc
OBJ := _build
LIB := cufhedb_amd/libcufhedb_amd.so

HIP_SRCS := $(wildcard $(SRC)/*.hip)
CPP_SRCS := $(wildcard $(SRC)/*.cpp)
HDRS := $(wildcard $(SRC)/*.hpp) include/cufhedb_amd.h
OBJS := $(patsubst $(SRC)/%.hip,$(OBJ)/%.hip.o,$(HIP_SRCS)) \
        $(patsubst $(SRC)/%.cpp,$(OBJ)/%.cpp.o,$(CPP_SRCS))

.PHONY: all lib oracle ref dropin clean
all: lib oracle dropin

lib: $(LIB)

$(OBJ)/%.hip.o: $(SRC)/%.hip $(HDRS)
	@mkdir -p $(OBJ)
	$(HIPCC) $(HIPFLAGS) -c $< -o $@

$(OBJ)/%.cpp.o: $(SRC)/%.cpp $(HDRS)
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(LIB): $(OBJS)
	$(HIPCC) --offload-arch=$(ARCH) -shared -fPIC -o $@ $(OBJS) -lpthread

# C++ drop-in usage examples / test programs (run by tests/test_gpu_dropin.py)
dropin: $(OBJ)/dropin_demo $(OBJ)/dropin_refcalls

$(OBJ)/dropin_refcalls: tests/cpp/dropin_refcalls.cpp include/cufhedb_amd/cuhedb.hpp include/cufhedb_amd.h $(LIB)
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -Iinclude -o $@ $< -Lcufhedb_amd -lcufhedb_amd -pthread -Wl,-rpath,'$$ORIGIN/../cufhedb_amd'

$(OBJ)/dropin_demo: tests/cpp/dropin_demo.cpp include/cufhedb_amd/cuhedb.hpp include/cufhedb_amd.h $(LIB)
	@mkdir -p $(OBJ)
	$(CXX) $(CXXFLAGS) -Iinclude -o $@ $< -Lcufhedb_amd -lcufhedb_amd -Wl,-rpath,'$$ORIGIN/../cufhedb_amd'

oracle:
	$(MAKE) -C oracle all

ref:
	$(MAKE) -C oracle ref

clean:
	rm -rf $(OBJ) $(LIB)
	$(MAKE) -C oracle clean
