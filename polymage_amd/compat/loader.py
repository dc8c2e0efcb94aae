"""``loader.load_lib`` (apps/python/loader.py:4-21): ctypes load of the
pipeline library, storing the function in app_data like the reference."""
import ctypes


def load_lib(so_file, func_name, app_data):
    lib = ctypes.cdll.LoadLibrary(so_file)
    app_data[func_name] = lib[func_name]
    app_data['lib'] = lib
    return
