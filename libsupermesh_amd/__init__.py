"""MI355X-native supermeshing engine with the libsupermesh C ABI."""
