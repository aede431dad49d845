"""MI355X path-tracing back end: HIP kernels behind the C ABI in
include/rtg.h (librtg.so), the C++ host side (JSON scenes, Scene::compile,
binned-SAH BVH, GpuRenderer, the `rt` CLI; librthost.so) and their ctypes
bindings. Import through the package loader in __graft_entry__.py."""
