"""MI355X-native PatchMatch multi-view stereo engine."""
