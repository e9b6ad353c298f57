"""The reference-side binding shown in INTEGRATION.md compiles.

The C++ `Force` subclass of INTEGRATION.md (its first C++ block) is compiled with g++
against this library's header and a minimal class of the same shape as rbmd-gpu's
`Force` (rbmd-gpu/src/force/include/force.h:8-25: a constructor, virtual Init and
Execute, protected `_structure_info_data` and `_device_data` shared pointers), with
the few DataManager / StructureInfoData / DeviceData members the subclass touches.
It is then linked against librbmd_force.so when that library is built, so every
entry point the binding calls must be exported.  No GPU call is made.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# The shape of rbmd-gpu's Force and of the members the binding reads (types as the
# reference declares them: rbmd::Real is float, device arrays are thrust vectors).
FORCE_H = r"""
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
namespace rbmd { using Real = float; using Id = int; }
namespace thrust {
template <class T> struct device_vector {
  std::vector<T> v;
  T* data() { return v.data(); }
};
template <class T> T* raw_pointer_cast(T* p) { return p; }
}
struct ConfigData {
  template <class T> T Get(const char*, const char*, const char*) const { return T(); }
};
struct StructureInfoData {
  rbmd::Id _num_atoms = 0, _num_atoms_type = 1;
  rbmd::Real _range[3][2] = {};
};
struct DeviceData {
  thrust::device_vector<rbmd::Real> _d_fx, _d_fy, _d_fz;
};
struct MDData { std::shared_ptr<StructureInfoData> _structure_info_data; };
class DataManager {
 public:
  static DataManager& getInstance() { static DataManager d; return d; }
  std::shared_ptr<ConfigData> getConfigData() { return _cfg; }
  std::shared_ptr<MDData> getMDData() { return _md; }
  std::shared_ptr<DeviceData> getDeviceData() { return _dev; }
 private:
  std::shared_ptr<ConfigData> _cfg = std::make_shared<ConfigData>();
  std::shared_ptr<MDData> _md = std::make_shared<MDData>();
  std::shared_ptr<DeviceData> _dev = std::make_shared<DeviceData>();
};
class Force {
 public:
  Force() {
    this->_structure_info_data = DataManager::getInstance().getMDData()->_structure_info_data;
    this->_device_data = DataManager::getInstance().getDeviceData();
  }
  virtual ~Force() = default;
  virtual void Init() {}
  virtual void Execute() = 0;
 protected:
  std::shared_ptr<StructureInfoData> _structure_info_data;
  std::shared_ptr<DeviceData> _device_data;
};
"""

# Constructs the subclass and calls both entry points (the program is linked, never
# run), so every library call in the binding is emitted and must resolve.
MAIN_CPP = r"""
#include "mi355_charged_lj_force.h"
int main(int argc, char**) {
  if (argc > 1000) {
    MI355ChargedLJForce f;
    Force& base = f;
    base.Init();
    base.Execute();
  }
  return 0;
}
"""


def _binding_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text.index("## rbmd-gpu: a `Force` subclass")
    m = re.search(r"```cpp\n(.*?)```", text[sec:], re.S)
    assert m, "no C++ block in the Force subclass section"
    return m.group(1)


def test_binding_block_names_the_interface():
    src = _binding_source()
    assert "class MI355ChargedLJForce : public Force" in src
    for fn in ("rbmd_force_create", "rbmd_force_set_atoms", "rbmd_force_sample_kvectors",
               "rbmd_force_compute", "rbmd_force_destroy", "rbmd_force_last_error"):
        assert fn + "(" in src, fn


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_binding_compiles_and_links(tmp_path):
    (tmp_path / "force.h").write_text(FORCE_H)
    (tmp_path / "mi355_charged_lj_force.h").write_text(_binding_source())
    (tmp_path / "main.cpp").write_text(MAIN_CPP)
    inc = ["-I", str(tmp_path), "-I", os.path.join(ROOT, "include")]
    obj = tmp_path / "main.o"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-Wno-unused-private-field",
                        *inc, "-c", str(tmp_path / "main.cpp"), "-o", str(obj)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = os.path.join(ROOT, "copy_gpu_amd", "librbmd_force.so")
    if not os.path.exists(lib):
        pytest.skip("library not built: compile-only check done")
    exe = tmp_path / "main"
    r = subprocess.run(["g++", str(obj), "-o", str(exe), lib, "-Wl,--allow-shlib-undefined",
                        f"-Wl,-rpath,{os.path.dirname(lib)}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
