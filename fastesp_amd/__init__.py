"""fastesp_amd: MI355X-native backend for fastesp's ESP hot path."""
