"""MI355X-native CLEO super-droplet hot path: ctypes access to libcleo_hip.so."""
