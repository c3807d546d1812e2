"""include/rz.h is a valid C99 and C++17 header: a C consumer compiles
against it, checks the layout of rz_batch_item that foreign bindings (the
numpy record in rusterizer_amd/_lib.py, the Rust #[repr(C)] struct) rely on,
links against librz.so and calls the entry points that need no device."""
import os
import shutil
import subprocess

import pytest

from rusterizer_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CONSUMER = r"""
#include <stdio.h>
#include <string.h>
#include <stddef.h>
#include "rz.h"

#ifdef __cplusplus
#define STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

/* the layout every foreign binding of rz_batch_item assumes */
STATIC_ASSERT(sizeof(rz_batch_item) == 232, "rz_batch_item size");
STATIC_ASSERT(offsetof(rz_batch_item, width) == 0, "width");
STATIC_ASSERT(offsetof(rz_batch_item, height) == 4, "height");
STATIC_ASSERT(offsetof(rz_batch_item, clear_depth) == 8, "clear_depth");
STATIC_ASSERT(offsetof(rz_batch_item, texture) == 12, "texture");
STATIC_ASSERT(offsetof(rz_batch_item, vertex_shader) == 16, "vertex_shader");
STATIC_ASSERT(offsetof(rz_batch_item, fragment_shader) == 20, "fragment_shader");
STATIC_ASSERT(offsetof(rz_batch_item, argb_dev) == 24, "argb_dev");
STATIC_ASSERT(offsetof(rz_batch_item, mesh) == 32, "mesh");
STATIC_ASSERT(offsetof(rz_batch_item, world) == 40, "world");
STATIC_ASSERT(offsetof(rz_batch_item, view) == 104, "view");
STATIC_ASSERT(offsetof(rz_batch_item, proj) == 168, "proj");
STATIC_ASSERT(RZ_OK == 0 && RZ_ERR_ARG == -1 && RZ_ERR_HIP == -2 && RZ_ERR_STATE == -3 &&
              RZ_ERR_NOMEM == -4 && RZ_ERR_LIMIT == -5, "status codes");
STATIC_ASSERT(RZ_FS_TEXTURE == 0 && RZ_FS_COLOR == 1 && RZ_FS_DEBUG == 2 && RZ_VS_MVP == 0, "shader enums");
STATIC_ASSERT(RZ_NSTAGES == 3, "stages");

int main(void) {
    rz_ctx *ctx = NULL;
    uint64_t stats[11];
    if (strcmp(rz_last_error(NULL), "null context") != 0) return 1;
    if (rz_create(0, 10, &ctx) != RZ_ERR_ARG || ctx) return 2;
    if (rz_create(16384, 16384, &ctx) != RZ_ERR_LIMIT || ctx) return 3;
    if (rz_destroy(NULL) != RZ_OK) return 4;
    if (rz_draw(NULL, NULL, RZ_VS_MVP, RZ_FS_COLOR) != RZ_ERR_ARG) return 5;
    if (rz_frame_stats(NULL, stats) != RZ_ERR_ARG) return 6;
    if (rz_batch_add(NULL, NULL, 0, NULL) != RZ_ERR_ARG) return 7;
    if (rz_texture_count(NULL) != 0 || rz_frame_flags(NULL) != 0 || rz_stream(NULL) != NULL) return 8;
    printf("ok\n");
    return 0;
}
"""


def _compile_and_run(tmp_path, compiler, std, suffix):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not found")
    src = tmp_path / f"consumer.{suffix}"
    src.write_text(CONSUMER)
    exe = tmp_path / f"consumer_{suffix}"
    libdir = os.path.dirname(_lib.lib_path())
    cmd = [compiler, f"-std={std}", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           str(src), "-o", str(exe), f"-L{libdir}", "-lrz", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout, r.stderr)


def test_header_compiles_as_c99_and_runs(tmp_path):
    _compile_and_run(tmp_path, "gcc", "c99", "c")


def test_header_compiles_as_cpp17_and_runs(tmp_path):
    _compile_and_run(tmp_path, "g++", "c++17", "cpp")
