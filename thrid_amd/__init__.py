"""MI355X-native force-evaluation hot path of peachgk_md (calforce / calforce_hf)."""
